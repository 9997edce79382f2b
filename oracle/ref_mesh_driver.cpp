// C entry point around the reference's binary mesh writer: a triangle mesh
// in sprout's JSON mesh format, parsed by its own SAX handler
// (serialize::Json_handler, triangle_json_handler.cpp), written by its own
// serialize::Exporter::write (triangle_mesh_exporter.cpp) as
// <name>.sub in the current directory.  Test infrastructure only: it makes
// the binary mesh fixtures the take loader is checked against.

#include "scene/shape/triangle/triangle_json_handler.hpp"
#include "scene/shape/triangle/triangle_mesh_exporter.hpp"

#include "rapidjson/reader.h"

#include <string>

extern "C" int sref_export_mesh(char const* json_text, char const* name) {
    scene::shape::triangle::serialize::Json_handler handler;

    rapidjson::Reader       reader;
    rapidjson::StringStream stream(json_text);
    if (!reader.Parse(stream, handler)) {
        return -1;
    }
    scene::shape::triangle::serialize::Exporter::write(std::string(name), handler);
    return 0;
}
