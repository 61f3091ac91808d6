// Logger.h -- the reference's Logger(level) << ... lines (Logger.h), minimal.
#pragma once
#include <iostream>
#include <sstream>

namespace mlh_host {

enum LogLevel { DEBUG = 0, INFO = 1, WARN = 2, ERROR = 3 };

struct LogConfig {
    bool headers = true;
    LogLevel level = INFO;
};
LogConfig &log_config();

class Logger {
public:
    explicit Logger(LogLevel l) : level_(l) {
        if (log_config().headers) {
            static const char *names[] = {"[DEBUG] ", "[INFO] ", "[WARN] ", "[ERROR] "};
            os_ << names[l];
        }
    }
    ~Logger() {
        if (level_ >= log_config().level) (level_ >= WARN ? std::cerr : std::cout) << os_.str() << "\n";
    }
    template <typename T>
    Logger &operator<<(const T &v) {
        os_ << v;
        return *this;
    }

private:
    LogLevel level_;
    std::ostringstream os_;
};

}  // namespace mlh_host
