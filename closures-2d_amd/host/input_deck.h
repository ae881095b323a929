// input_deck.h -- key/value input deck, with the semantics of the reference's
// InputDeckReader (src/input_deck_reader.{h,cpp}): one "KEY VALUE" pair per
// line, '#' starts a comment, keys and values are matched case-insensitively
// (values are lower-cased), the first occurrence of a key wins, and a missing
// key is reported to the caller.
#pragma once

#include <string>
#include <utility>
#include <vector>

namespace mnc_host {

class InputDeck {
public:
    // Returns false if the file cannot be opened.
    bool read(const std::string &path);
    bool parse_text(const std::string &text);
    bool get(const std::string &key, std::string &value) const;
    bool get(const std::string &key, int &value) const;
    bool get(const std::string &key, double &value) const;
    void print() const;

private:
    std::vector<std::pair<std::string, std::string>> kv_;
};

}  // namespace mnc_host
