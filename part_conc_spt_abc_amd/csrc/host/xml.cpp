// xml.cpp -- recursive-descent reader for the flat XML files of the
// input/data/output contract.
#include "xml.hpp"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace pcs::xml {

const Element* Element::child(const std::string& n) const
{
    for (const auto& c : children)
        if (c->name == n) return c.get();
    return nullptr;
}

namespace {

class Reader {
public:
    explicit Reader(const std::string& s) : s_(s) {}

    Document document()
    {
        Document d;
        skip_misc();
        if (pos_ >= s_.size() || s_[pos_] != '<') fail("expected a root element");
        d.root = element();
        skip_misc();
        if (pos_ != s_.size()) fail("content after the root element");
        return d;
    }

private:
    const std::string& s_;
    size_t pos_ = 0;

    [[noreturn]] void fail(const char* what) const
    {
        size_t line = 1;
        for (size_t i = 0; i < pos_ && i < s_.size(); ++i) line += s_[i] == '\n';
        throw ParseError(std::string("XML: ") + what + " (line " + std::to_string(line) + ")");
    }

    bool starts(const char* p) const { return s_.compare(pos_, std::char_traits<char>::length(p), p) == 0; }

    void skip_ws()
    {
        while (pos_ < s_.size() && (s_[pos_] == ' ' || s_[pos_] == '\t' || s_[pos_] == '\n' || s_[pos_] == '\r'))
            ++pos_;
    }

    void skip_until(const char* end)
    {
        const size_t e = s_.find(end, pos_);
        if (e == std::string::npos) fail("unterminated markup");
        pos_ = e + std::char_traits<char>::length(end);
    }

    // whitespace, comments, processing instructions, doctype
    void skip_misc()
    {
        for (;;) {
            skip_ws();
            if (starts("<?")) skip_until("?>");
            else if (starts("<!--")) skip_until("-->");
            else if (starts("<!DOCTYPE")) skip_until(">");
            else return;
        }
    }

    std::string name()
    {
        const size_t b = pos_;
        while (pos_ < s_.size()) {
            const char c = s_[pos_];
            if (c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '>' || c == '/' || c == '=') break;
            ++pos_;
        }
        if (pos_ == b) fail("expected a name");
        return s_.substr(b, pos_ - b);
    }

    void append_decoded(std::string& out, size_t b, size_t e)
    {
        for (size_t i = b; i < e; ++i) {
            if (s_[i] != '&') {
                out.push_back(s_[i]);
                continue;
            }
            const size_t semi = s_.find(';', i);
            if (semi == std::string::npos || semi > e) fail("unterminated entity");
            const std::string ent = s_.substr(i + 1, semi - i - 1);
            if (ent == "lt") out.push_back('<');
            else if (ent == "gt") out.push_back('>');
            else if (ent == "amp") out.push_back('&');
            else if (ent == "quot") out.push_back('"');
            else if (ent == "apos") out.push_back('\'');
            else if (!ent.empty() && ent[0] == '#') {
                const long cp = ent.size() > 1 && (ent[1] == 'x' || ent[1] == 'X') ? std::strtol(ent.c_str() + 2, nullptr, 16)
                                                                                  : std::strtol(ent.c_str() + 1, nullptr, 10);
                if (cp <= 0 || cp > 0x10FFFF) fail("bad character reference");
                // UTF-8 encode
                if (cp < 0x80) out.push_back((char)cp);
                else if (cp < 0x800) {
                    out.push_back((char)(0xC0 | (cp >> 6)));
                    out.push_back((char)(0x80 | (cp & 0x3F)));
                } else if (cp < 0x10000) {
                    out.push_back((char)(0xE0 | (cp >> 12)));
                    out.push_back((char)(0x80 | ((cp >> 6) & 0x3F)));
                    out.push_back((char)(0x80 | (cp & 0x3F)));
                } else {
                    out.push_back((char)(0xF0 | (cp >> 18)));
                    out.push_back((char)(0x80 | ((cp >> 12) & 0x3F)));
                    out.push_back((char)(0x80 | ((cp >> 6) & 0x3F)));
                    out.push_back((char)(0x80 | (cp & 0x3F)));
                }
            } else {
                fail("unknown entity");
            }
            i = semi;
        }
    }

    std::unique_ptr<Element> element()
    {
        ++pos_;  // '<'
        auto el = std::make_unique<Element>();
        el->name = name();
        for (;;) {
            skip_ws();
            if (pos_ >= s_.size()) fail("unterminated start tag");
            if (starts("/>")) {
                pos_ += 2;
                return el;
            }
            if (s_[pos_] == '>') {
                ++pos_;
                break;
            }
            std::string an = name();
            skip_ws();
            if (pos_ >= s_.size() || s_[pos_] != '=') fail("expected '=' after attribute name");
            ++pos_;
            skip_ws();
            if (pos_ >= s_.size() || (s_[pos_] != '"' && s_[pos_] != '\'')) fail("expected quoted attribute value");
            const char q = s_[pos_++];
            const size_t e = s_.find(q, pos_);
            if (e == std::string::npos) fail("unterminated attribute value");
            std::string av;
            append_decoded(av, pos_, e);
            pos_ = e + 1;
            el->attributes.emplace_back(std::move(an), std::move(av));
        }
        // content
        for (;;) {
            if (pos_ >= s_.size()) fail("unterminated element");
            if (starts("</")) {
                pos_ += 2;
                const std::string n = name();
                if (n != el->name) fail("mismatched end tag");
                skip_ws();
                if (pos_ >= s_.size() || s_[pos_] != '>') fail("expected '>'");
                ++pos_;
                return el;
            }
            if (starts("<!--")) {
                skip_until("-->");
                continue;
            }
            if (starts("<![CDATA[")) {
                const size_t b = pos_ + 9;
                const size_t e = s_.find("]]>", b);
                if (e == std::string::npos) fail("unterminated CDATA");
                el->text.append(s_, b, e - b);
                pos_ = e + 3;
                continue;
            }
            if (starts("<?")) {
                skip_until("?>");
                continue;
            }
            if (s_[pos_] == '<') {
                el->children.push_back(element());
                continue;
            }
            const size_t e = s_.find('<', pos_);
            if (e == std::string::npos) fail("unterminated element");
            append_decoded(el->text, pos_, e);
            pos_ = e;
        }
    }
};

}  // namespace

Document parse(const std::string& text) { return Reader(text).document(); }

Document load_file(const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open '" + path + "'");
    std::ostringstream ss;
    ss << f.rdbuf();
    return parse(ss.str());
}

}  // namespace pcs::xml
