// ConfigParser.cpp -- INFO-format reader (see ConfigParser.h).
#include "ConfigParser.h"

#include <cctype>
#include <fstream>

namespace mlh_host {

namespace {

std::string trim(const std::string &s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

// strip a ';' comment that is not inside double quotes
std::string strip_comment(const std::string &s) {
    bool q = false;
    for (size_t k = 0; k < s.size(); ++k) {
        if (s[k] == '"') q = !q;
        if (s[k] == ';' && !q) return s.substr(0, k);
    }
    return s;
}

std::string unquote(const std::string &s) {
    if (s.size() >= 2 && s.front() == '"' && s.back() == '"') return s.substr(1, s.size() - 2);
    return s;
}

}  // namespace

ConfigParser::ConfigParser(const std::string &file) : root_(std::make_shared<Node>()) {
    std::ifstream in(file);
    if (!in) throw std::runtime_error("config: cannot open '" + file + "'");
    int line = 0;
    parse(in, *root_, 0, line);
}

void ConfigParser::parse(std::istream &in, Node &into, int depth, int &line) {
    std::string s;
    while (std::getline(in, s)) {
        ++line;
        std::string t = trim(strip_comment(s));
        if (t.empty()) continue;
        if (t == "}") {
            if (depth == 0) throw std::runtime_error("config: unmatched '}' at line " + std::to_string(line));
            return;
        }
        if (t == "{") {  // brace on its own line opens the previous key
            if (into.children.empty())
                throw std::runtime_error("config: '{' without a key at line " + std::to_string(line));
            parse(in, *into.children.back().second, depth + 1, line);
            continue;
        }
        size_t k = 0;
        while (k < t.size() && !std::isspace((unsigned char)t[k])) ++k;
        std::string key = t.substr(0, k);
        std::string rest = trim(t.substr(k));
        bool opens = false;
        if (!rest.empty() && rest.back() == '{') {
            opens = true;
            rest = trim(rest.substr(0, rest.size() - 1));
        }
        if (key == "{" || key.empty())
            throw std::runtime_error("config: malformed line " + std::to_string(line));
        auto n = std::make_shared<Node>();
        n->value = unquote(rest);
        into.children.emplace_back(key, n);
        if (opens) parse(in, *n, depth + 1, line);
    }
    if (depth != 0) throw std::runtime_error("config: missing '}' at end of file");
}

const ConfigParser::Node &ConfigParser::node(const std::string &key) const {
    const Node *n = root_.get();
    size_t a = 0;
    while (a <= key.size()) {
        size_t b = key.find('.', a);
        if (b == std::string::npos) b = key.size();
        std::string part = key.substr(a, b - a);
        const Node *next = nullptr;
        for (const auto &c : n->children)
            if (c.first == part) { next = c.second.get(); break; }  // first match, as ptree::get
        if (!next) throw std::runtime_error("config: no key '" + key + "'");
        n = next;
        a = b + 1;
    }
    return *n;
}

std::string ConfigParser::raw(const std::string &key) const { return node(key).value; }

bool ConfigParser::has(const std::string &key) const {
    try {
        node(key);
        return true;
    } catch (const std::runtime_error &) {
        return false;
    }
}

ConfigParser ConfigParser::getObj(const std::string &key) const {
    const Node &n = node(key);
    // share the sub-tree (read-only use)
    auto copy = std::make_shared<Node>(n);
    return ConfigParser(copy);
}

std::vector<std::pair<std::string, std::string>> ConfigParser::flatten() const {
    std::vector<std::pair<std::string, std::string>> out;
    struct Walk {
        static void go(const Node &n, const std::string &pre,
                       std::vector<std::pair<std::string, std::string>> &o) {
            for (const auto &c : n.children) {
                std::string p = pre.empty() ? c.first : pre + "." + c.first;
                if (c.second->children.empty()) o.emplace_back(p, c.second->value);
                else go(*c.second, p, o);
            }
        }
    };
    Walk::go(*root_, "", out);
    return out;
}

}  // namespace mlh_host
