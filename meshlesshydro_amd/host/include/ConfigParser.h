// ConfigParser.h -- reader of the reference's config.info files.
//
// Mirrors the interface of ConfigParser (demonstrator/include/ConfigParser.h:
// getVal<T>(key), getObj(key), getList<T>(key)) without Boost.  The file
// format is the INFO format of Boost.PropertyTree as the reference's configs
// use it: one "key value" per line, ';' starts a comment, "key { ... }"
// opens a sub-tree, values may be double-quoted.
#pragma once
#include <list>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace mlh_host {

class ConfigParser {
public:
    explicit ConfigParser(const std::string &file = "config.info");

    // sub-tree of a "key { ... }" block (ConfigParser.cpp getObj)
    ConfigParser getObj(const std::string &key) const;

    // value of "key value"; a '.' in key descends into sub-trees
    template <typename T>
    T getVal(const std::string &key) const {
        std::istringstream in(raw(key));
        T v;
        in >> v;
        if (in.fail()) throw std::runtime_error("config: cannot convert value of '" + key + "'");
        return v;
    }

    // the values of a block's children, in file order
    template <typename T>
    std::list<T> getList(const std::string &key) const {
        std::list<T> out;
        const Node &n = node(key);
        for (const auto &c : n.children) {
            std::istringstream in(c.second->value);
            T v;
            in >> v;
            if (in.fail()) throw std::runtime_error("config: list '" + key + "' holds a non-value");
            out.push_back(v);
        }
        return out;
    }

    bool has(const std::string &key) const;
    // all "path = value" pairs, for diagnostics
    std::vector<std::pair<std::string, std::string>> flatten() const;

private:
    struct Node {
        std::string value;
        std::vector<std::pair<std::string, std::shared_ptr<Node>>> children;
    };
    explicit ConfigParser(std::shared_ptr<Node> root) : root_(std::move(root)) {}
    const Node &node(const std::string &key) const;
    std::string raw(const std::string &key) const;
    static void parse(std::istream &in, Node &into, int depth, int &line);
    std::shared_ptr<Node> root_;
};

template <>
inline std::string ConfigParser::getVal<std::string>(const std::string &key) const {
    return raw(key);
}

}  // namespace mlh_host
