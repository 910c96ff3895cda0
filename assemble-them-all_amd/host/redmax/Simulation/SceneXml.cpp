#include "Simulation/SceneXml.h"

#include <cctype>
#include <cstdlib>

#include "Common.h"

namespace redmax {
namespace xml {

const Node* Node::child(const std::string& n) const {
    for (const auto& c : children)
        if (c.name == n) return &c;
    return nullptr;
}

const std::string* Node::attr(const std::string& n) const {
    for (const auto& a : attrs)
        if (a.first == n) return &a.second;
    return nullptr;
}

namespace {

class Reader {
public:
    explicit Reader(const std::string& s) : s_(s) {}

    Node document() {
        Node doc;
        while (true) {
            skip_misc();
            if (at_end()) break;
            if (s_[i_] != '<') fail("text outside the root element");
            doc.children.push_back(element());
        }
        return doc;
    }

private:
    const std::string& s_;
    size_t i_ = 0;

    [[noreturn]] void fail(const std::string& what) const {
        size_t line = 1;
        for (size_t k = 0; k < i_ && k < s_.size(); ++k) line += s_[k] == '\n';
        throw_error("scene XML: " + what + " (line " + std::to_string(line) + ")");
        throw what;   // not reached: throw_error throws
    }
    bool at_end() const { return i_ >= s_.size(); }
    bool starts(const char* p) const { return s_.compare(i_, std::char_traits<char>::length(p), p) == 0; }
    void skip_ws() {
        while (!at_end() && std::isspace((unsigned char)s_[i_])) ++i_;
    }
    void skip_past(const char* end) {
        const size_t k = s_.find(end, i_);
        if (k == std::string::npos) fail(std::string("unterminated construct, missing '") + end + "'");
        i_ = k + std::char_traits<char>::length(end);
    }
    // whitespace, comments, processing instructions, DOCTYPE
    void skip_misc() {
        while (true) {
            skip_ws();
            if (starts("<!--")) skip_past("-->");
            else if (starts("<?")) skip_past("?>");
            else if (starts("<!DOCTYPE") || starts("<!doctype")) skip_past(">");
            else break;
        }
    }
    static bool name_char(char c) {
        return std::isalnum((unsigned char)c) || c == '_' || c == '-' || c == ':' || c == '.';
    }
    std::string name() {
        const size_t b = i_;
        while (!at_end() && name_char(s_[i_])) ++i_;
        if (i_ == b) fail("expected a name");
        return s_.substr(b, i_ - b);
    }
    std::string decode(const std::string& raw) const {
        std::string out;
        for (size_t k = 0; k < raw.size(); ++k) {
            if (raw[k] != '&') {
                out += raw[k];
                continue;
            }
            const size_t e = raw.find(';', k);
            if (e == std::string::npos) {
                out += raw[k];
                continue;
            }
            const std::string ent = raw.substr(k + 1, e - k - 1);
            if (ent == "lt") out += '<';
            else if (ent == "gt") out += '>';
            else if (ent == "amp") out += '&';
            else if (ent == "quot") out += '"';
            else if (ent == "apos") out += '\'';
            else if (ent.size() > 1 && ent[0] == '#') {
                const long cp = ent[1] == 'x' ? std::strtol(ent.c_str() + 2, nullptr, 16) : std::strtol(ent.c_str() + 1, nullptr, 10);
                if (cp > 0 && cp < 128) out += (char)cp;   // ASCII only: scene attributes are numbers and names
            } else {
                out += raw.substr(k, e - k + 1);
            }
            k = e;
        }
        return out;
    }
    Node element() {
        ++i_;   // '<'
        Node n;
        n.name = name();
        while (true) {
            skip_ws();
            if (at_end()) fail("unterminated start tag <" + n.name + ">");
            if (starts("/>")) {
                i_ += 2;
                return n;
            }
            if (s_[i_] == '>') {
                ++i_;
                break;
            }
            std::string key = name();
            skip_ws();
            if (at_end() || s_[i_] != '=') fail("attribute " + key + " without a value");
            ++i_;
            skip_ws();
            if (at_end() || (s_[i_] != '"' && s_[i_] != '\'')) fail("unquoted value of attribute " + key);
            const char q = s_[i_++];
            const size_t e = s_.find(q, i_);
            if (e == std::string::npos) fail("unterminated value of attribute " + key);
            n.attrs.emplace_back(std::move(key), decode(s_.substr(i_, e - i_)));
            i_ = e + 1;
        }
        // content: child elements; text, comments, CDATA skipped
        while (true) {
            if (at_end()) fail("missing </" + n.name + ">");
            if (starts("<!--")) skip_past("-->");
            else if (starts("<![CDATA[")) skip_past("]]>");
            else if (starts("<?")) skip_past("?>");
            else if (starts("</")) {
                i_ += 2;
                const std::string end = name();
                if (end != n.name) fail("</" + end + "> closes <" + n.name + ">");
                skip_ws();
                if (at_end() || s_[i_] != '>') fail("malformed end tag </" + end);
                ++i_;
                return n;
            } else if (s_[i_] == '<') {
                n.children.push_back(element());
            } else {
                ++i_;   // text
            }
        }
    }
};

}  // namespace

Node parse(const std::string& text) { return Reader(text).document(); }

}  // namespace xml
}  // namespace redmax
