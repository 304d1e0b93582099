// Minimal XML reader for ARTSS input files (elements, attributes, text,
// comments, processing instructions, the five predefined entities).
// The reference reads the same files with tinyxml2 (utility/settings/tinyxml2.h);
// only the document tree is needed here, so this is a small recursive-descent
// parser of its own.
#pragma once

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace artss {
namespace xml {

struct Element {
    std::string name;
    std::vector<std::pair<std::string, std::string>> attributes;  // document order
    std::string text;                                             // first text run, as in tinyxml2 GetText()
    bool has_text = false;
    std::vector<std::unique_ptr<Element>> children;

    const Element *child(const std::string &n) const;
    const std::string *attribute(const std::string &n) const;
};

class parse_error : public std::runtime_error {
  public:
    explicit parse_error(const std::string &m) : std::runtime_error(m) {}
};

std::unique_ptr<Element> parse(const std::string &text);

}  // namespace xml
}  // namespace artss
