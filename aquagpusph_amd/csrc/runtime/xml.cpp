#include "xml.hpp"

#include <cctype>
#include <cstring>
#include <fstream>
#include <sstream>

#include "../aq_internal.hpp"

namespace aq {
namespace rt {

bool XmlNode::has(const std::string& a) const
{
    for (auto& kv : attrs)
        if (kv.first == a) return true;
    return false;
}

std::string XmlNode::attr(const std::string& a, const std::string& def) const
{
    for (auto& kv : attrs)
        if (kv.first == a) return kv.second;
    return def;
}

void XmlNode::find_all(const std::string& tag,
                       std::vector<const XmlNode*>& out) const
{
    for (auto& c : children) {
        if (c->name == tag) out.push_back(c.get());
        c->find_all(tag, out);
    }
}

std::string XmlNode::text_content() const
{
    std::string s = text;
    for (auto& c : children) s += c->text_content();
    return s;
}

namespace {

struct Parser {
    const std::string& s;
    const std::string& origin;
    size_t i = 0;
    int line = 1;

    [[noreturn]] void fail(const std::string& m)
    {
        std::ostringstream o;
        o << origin << ":" << line << ": XML error: " << m;
        throw Error(AQ_EPARSE, o.str());
    }
    bool eof() const { return i >= s.size(); }
    char peek(size_t k = 0) const { return i + k < s.size() ? s[i + k] : '\0'; }
    void adv(size_t n = 1)
    {
        for (size_t k = 0; k < n && i < s.size(); k++, i++)
            if (s[i] == '\n') line++;
    }
    bool starts(const char* p) const { return s.compare(i, strlen(p), p) == 0; }
    void skip_ws()
    {
        while (!eof() && isspace((unsigned char)peek())) adv();
    }
    void skip_until(const char* end)
    {
        const size_t n = strlen(end);
        while (!eof() && !starts(end)) adv();
        if (eof()) fail(std::string("unterminated construct, expected '") + end + "'");
        adv(n);
    }
    std::string name()
    {
        const size_t b = i;
        while (!eof() && (isalnum((unsigned char)peek()) || peek() == '_' ||
                          peek() == '-' || peek() == '.' || peek() == ':'))
            adv();
        if (i == b) fail("expected a name");
        return s.substr(b, i - b);
    }
    std::string decode(const std::string& raw)
    {
        std::string o;
        for (size_t k = 0; k < raw.size(); k++) {
            if (raw[k] != '&') {
                o += raw[k];
                continue;
            }
            const size_t e = raw.find(';', k);
            if (e == std::string::npos) fail("unterminated entity");
            const std::string ent = raw.substr(k + 1, e - k - 1);
            if (ent == "lt") o += '<';
            else if (ent == "gt") o += '>';
            else if (ent == "amp") o += '&';
            else if (ent == "quot") o += '"';
            else if (ent == "apos") o += '\'';
            else if (!ent.empty() && ent[0] == '#') {
                unsigned long c = (ent.size() > 1 && ent[1] == 'x')
                                      ? strtoul(ent.c_str() + 2, nullptr, 16)
                                      : strtoul(ent.c_str() + 1, nullptr, 10);
                if (c < 0x80) o += (char)c;
                else fail("non-ASCII character reference");
            } else fail("unknown entity &" + ent + ";");
            k = e;
        }
        return o;
    }
    // skip prolog items, comments, PIs, doctype, whitespace
    void misc()
    {
        for (;;) {
            skip_ws();
            if (starts("<?")) skip_until("?>");
            else if (starts("<!--")) skip_until("-->");
            else if (starts("<!DOCTYPE")) skip_until(">");
            else return;
        }
    }
    std::unique_ptr<XmlNode> element()
    {
        if (peek() != '<') fail("expected '<'");
        adv();
        auto n = std::make_unique<XmlNode>();
        n->line = line;
        n->name = name();
        for (;;) {
            skip_ws();
            if (starts("/>")) {
                adv(2);
                return n;
            }
            if (peek() == '>') {
                adv();
                break;
            }
            std::string an = name();
            skip_ws();
            if (peek() != '=') fail("expected '=' after attribute " + an);
            adv();
            skip_ws();
            const char q = peek();
            if (q != '"' && q != '\'') fail("expected quoted attribute value");
            adv();
            const size_t b = i;
            while (!eof() && peek() != q) adv();
            if (eof()) fail("unterminated attribute value");
            n->attrs.emplace_back(an, decode(s.substr(b, i - b)));
            adv();
        }
        // content
        for (;;) {
            if (eof()) fail("unterminated element <" + n->name + ">");
            if (starts("</")) {
                adv(2);
                const std::string en = name();
                if (en != n->name)
                    fail("mismatched </" + en + ">, expected </" + n->name + ">");
                skip_ws();
                if (peek() != '>') fail("expected '>'");
                adv();
                return n;
            }
            if (starts("<!--")) {
                skip_until("-->");
            } else if (starts("<![CDATA[")) {
                adv(9);
                const size_t b = i;
                while (!eof() && !starts("]]>")) adv();
                if (eof()) fail("unterminated CDATA");
                n->text += s.substr(b, i - b);
                adv(3);
            } else if (starts("<?")) {
                skip_until("?>");
            } else if (peek() == '<') {
                n->children.push_back(element());
            } else {
                const size_t b = i;
                while (!eof() && peek() != '<') adv();
                n->text += decode(s.substr(b, i - b));
            }
        }
    }
};

}  // namespace

std::unique_ptr<XmlNode> xml_parse(const std::string& text,
                                   const std::string& origin)
{
    Parser p{ text, origin };
    p.misc();
    if (p.eof()) p.fail("empty document");
    auto root = p.element();
    p.misc();
    if (!p.eof()) p.fail("content after the root element");
    return root;
}

std::unique_ptr<XmlNode> xml_parse_file(const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error(AQ_ENOTFOUND, "cannot open XML file '" + path + "'");
    std::stringstream ss;
    ss << f.rdbuf();
    return xml_parse(ss.str(), path);
}

}  // namespace rt
}  // namespace aq
