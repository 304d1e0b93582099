#include "xml.h"

#include <cctype>

namespace artss {
namespace xml {

const Element *Element::child(const std::string &n) const {
    for (const auto &c : children)
        if (c->name == n) return c.get();
    return nullptr;
}

const std::string *Element::attribute(const std::string &n) const {
    for (const auto &a : attributes)
        if (a.first == n) return &a.second;
    return nullptr;
}

namespace {

struct Parser {
    const std::string &s;
    size_t p = 0;
    explicit Parser(const std::string &text) : s(text) {}

    [[noreturn]] void fail(const std::string &what) const {
        throw parse_error("XML: " + what + " at offset " + std::to_string(p));
    }
    bool starts(const char *lit) const { return s.compare(p, std::char_traits<char>::length(lit), lit) == 0; }
    void skip_ws() {
        while (p < s.size() && std::isspace(static_cast<unsigned char>(s[p]))) p++;
    }
    void skip_until(const char *lit) {
        size_t e = s.find(lit, p);
        if (e == std::string::npos) fail(std::string("missing ") + lit);
        p = e + std::char_traits<char>::length(lit);
    }
    // comments, processing instructions, doctype
    bool skip_misc() {
        if (starts("<!--")) { p += 4; skip_until("-->"); return true; }
        if (starts("<?")) { p += 2; skip_until("?>"); return true; }
        if (starts("<!")) { p += 2; skip_until(">"); return true; }
        return false;
    }
    static bool name_char(char c) {
        return std::isalnum(static_cast<unsigned char>(c)) || c == '_' || c == '-' || c == '.' || c == ':';
    }
    std::string name() {
        size_t b = p;
        while (p < s.size() && name_char(s[p])) p++;
        if (b == p) fail("expected a name");
        return s.substr(b, p - b);
    }
    static std::string decode(const std::string &raw) {
        std::string out;
        out.reserve(raw.size());
        for (size_t i = 0; i < raw.size(); i++) {
            if (raw[i] != '&') { out += raw[i]; continue; }
            size_t e = raw.find(';', i);
            if (e == std::string::npos) { out += raw[i]; continue; }
            std::string ent = raw.substr(i + 1, e - i - 1);
            if (ent == "lt") out += '<';
            else if (ent == "gt") out += '>';
            else if (ent == "amp") out += '&';
            else if (ent == "quot") out += '"';
            else if (ent == "apos") out += '\'';
            else { out += raw.substr(i, e - i + 1); }
            i = e;
        }
        return out;
    }
    std::unique_ptr<Element> element() {
        if (s[p] != '<') fail("expected '<'");
        p++;
        auto el = std::make_unique<Element>();
        el->name = name();
        for (;;) {
            skip_ws();
            if (p >= s.size()) fail("unterminated tag");
            if (starts("/>")) { p += 2; return el; }
            if (s[p] == '>') { p++; break; }
            std::string an = name();
            skip_ws();
            if (p >= s.size() || s[p] != '=') fail("expected '='");
            p++;
            skip_ws();
            if (p >= s.size() || (s[p] != '"' && s[p] != '\'')) fail("expected quote");
            char q = s[p++];
            size_t e = s.find(q, p);
            if (e == std::string::npos) fail("unterminated attribute");
            el->attributes.emplace_back(an, decode(s.substr(p, e - p)));
            p = e + 1;
        }
        // content
        for (;;) {
            if (p >= s.size()) fail("unterminated element " + el->name);
            if (starts("</")) {
                p += 2;
                std::string n = name();
                if (n != el->name) fail("mismatched end tag " + n);
                skip_ws();
                if (p >= s.size() || s[p] != '>') fail("expected '>'");
                p++;
                return el;
            }
            if (skip_misc()) continue;
            if (s[p] == '<') {
                el->children.push_back(element());
                continue;
            }
            size_t e = s.find('<', p);
            if (e == std::string::npos) fail("unterminated text");
            std::string raw = s.substr(p, e - p);
            p = e;
            bool blank = true;
            for (char c : raw)
                if (!std::isspace(static_cast<unsigned char>(c))) blank = false;
            if (!blank && !el->has_text) {
                el->text = decode(raw);
                el->has_text = true;
            }
        }
    }
};

}  // namespace

std::unique_ptr<Element> parse(const std::string &text) {
    Parser ps(text);
    for (;;) {
        ps.skip_ws();
        if (ps.p >= text.size()) ps.fail("no root element");
        if (ps.skip_misc()) continue;
        break;
    }
    return ps.element();
}

}  // namespace xml
}  // namespace artss
