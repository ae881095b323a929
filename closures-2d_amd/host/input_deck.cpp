// input_deck.cpp -- see input_deck.h (behaviour of src/input_deck_reader.cpp:17-171).
#include "input_deck.h"

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace mnc_host {
namespace {
std::string lower(std::string s) {
    for (auto &c : s) c = char(std::tolower((unsigned char)c));
    return s;
}
}  // namespace

bool InputDeck::parse_text(const std::string &text) {
    std::istringstream in(text);
    std::string line;
    int lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        auto hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        std::istringstream ls(line);
        std::string key, value, extra;
        if (!(ls >> key)) continue;
        if (!(ls >> value)) {
            std::printf("Error in input file at line: %d\n", lineno);
            continue;
        }
        if (ls >> extra) {
            std::printf("Error in input file at line: %d\n", lineno);
            continue;
        }
        kv_.emplace_back(key, value);
    }
    return true;
}

bool InputDeck::read(const std::string &path) {
    std::ifstream f(path);
    if (!f) {
        std::printf("Could not open file.\n");
        return false;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    return parse_text(ss.str());
}

bool InputDeck::get(const std::string &key, std::string &value) const {
    const std::string k = lower(key);
    for (const auto &p : kv_) {
        if (lower(p.first) == k) {
            value = lower(p.second);
            return true;
        }
    }
    return false;
}

bool InputDeck::get(const std::string &key, int &value) const {
    std::string s;
    if (!get(key, s)) return false;
    value = std::atoi(s.c_str());
    return true;
}

bool InputDeck::get(const std::string &key, double &value) const {
    std::string s;
    if (!get(key, s)) return false;
    value = std::atof(s.c_str());
    return true;
}

void InputDeck::print() const {
    std::printf("--- Input Deck ---\n");
    for (const auto &p : kv_) std::printf("%s = %s\n", p.first.c_str(), p.second.c_str());
}

}  // namespace mnc_host
