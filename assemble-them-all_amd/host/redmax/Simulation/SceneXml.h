// A small XML reader for the simulator's scene descriptions (the input of
// Simulation(xml_string, asset_folder), simulation/redmax/Simulation_Constructor.cpp:123-141,
// which the reference reads with pugixml): elements with their attributes (single or
// double quoted, the five predefined entities and numeric character references
// decoded), nested or self-closing. Comments, processing instructions, the XML
// declaration, DOCTYPE and CDATA sections are skipped, and text content is ignored,
// since the scene format keeps everything in attributes.
#pragma once
#include <string>
#include <utility>
#include <vector>

namespace redmax {
namespace xml {

struct Node {
    std::string name;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::vector<Node> children;

    // first child element with this name (nullptr if none)
    const Node* child(const std::string& n) const;
    // attribute value (nullptr if absent)
    const std::string* attr(const std::string& n) const;
};

// The document: a node named "" whose children are the top-level elements. Throws a
// std::string (the simulator's throw_error convention) on malformed input.
Node parse(const std::string& text);

}  // namespace xml
}  // namespace redmax
