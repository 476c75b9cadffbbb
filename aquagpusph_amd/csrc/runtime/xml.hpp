// Minimal XML DOM reader for the case/preset files (the subset the
// reference parses with xerces-c in aquagpusph/InputOutput/State.cpp):
// elements, attributes, character data, comments, CDATA, the five predefined
// entities and numeric character references.  No DTD, namespaces or schema.
#pragma once
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace aq {
namespace rt {

struct XmlNode {
    std::string name;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::vector<std::unique_ptr<XmlNode>> children;
    std::string text;  // concatenated character data of this element
    int line = 0;

    bool has(const std::string& a) const;
    std::string attr(const std::string& a, const std::string& def = "") const;
    // all descendant elements named `tag`, in document order (like the DOM
    // getElementsByTagName the reference uses)
    void find_all(const std::string& tag, std::vector<const XmlNode*>& out) const;
    std::vector<const XmlNode*> all(const std::string& tag) const
    {
        std::vector<const XmlNode*> v;
        find_all(tag, v);
        return v;
    }
    // full text content of the element and its descendants
    std::string text_content() const;
};

// Parse a document; throws aq::Error(AQ_EPARSE) with file:line on errors.
std::unique_ptr<XmlNode> xml_parse(const std::string& text,
                                   const std::string& origin);
std::unique_ptr<XmlNode> xml_parse_file(const std::string& path);

}  // namespace rt
}  // namespace aq
