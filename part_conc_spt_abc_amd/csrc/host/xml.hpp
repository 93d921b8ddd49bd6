// xml.hpp -- a small XML reader for the flat documents of the input/data/
// output contract: elements, attributes (kept, unused), text, comments,
// processing instructions, the five predefined entities and numeric
// character references. Not a validating parser.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace pcs::xml {

struct Element {
    std::string name;
    std::string text;  // concatenated character data directly inside this element
    std::vector<std::pair<std::string, std::string>> attributes;
    std::vector<std::unique_ptr<Element>> children;

    // First child element with this name, or nullptr.
    const Element* child(const std::string& n) const;
};

struct Document {
    std::unique_ptr<Element> root;
};

class ParseError : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

Document parse(const std::string& text);
Document load_file(const std::string& path);  // throws ParseError / std::runtime_error

}  // namespace pcs::xml
