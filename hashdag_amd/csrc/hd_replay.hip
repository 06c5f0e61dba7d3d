// Replays: the reference's CSV action streams (ReplayManager, replay.h:12-635), parsed into
// hd_replay_action records with every edit scaled to the scene depth as replay.cpp applies it.
//
// Cells are split at ',' and lines at '\n'; the first cell names the action and trailing empty
// cells are dropped (ReplayManager::load_csv, replay.h:565-607). Numbers are read with strtof /
// strtol, as std::stof / std::stoi do: leading spaces are skipped and the number ends at the
// first character that cannot continue it; a cell with no number is an error (stof throws).
#include <hip/hip_runtime.h>

#include <cerrno>
#include <cstdlib>
#include <string>
#include <vector>

#include "hd_ctx.h"

struct hd_replay {
    std::vector<hd_replay_action> actions;
    uint32_t frames = 0;
};

namespace hd {
namespace {

struct ParseError {
    std::string msg;
};

float to_f(const std::string& cell) {
    const char* b = cell.c_str();
    char* e = nullptr;
    errno = 0;
    const float v = std::strtof(b, &e);
    if (e == b || errno == ERANGE) throw ParseError{"not a number: '" + cell + "'"};
    return v;
}

int to_i(const std::string& cell) {
    const char* b = cell.c_str();
    char* e = nullptr;
    errno = 0;
    const long v = std::strtol(b, &e, 10);
    if (e == b || errno == ERANGE || v < INT32_MIN || v > INT32_MAX) throw ParseError{"not an integer: '" + cell + "'"};
    return (int)v;
}

bool to_bool(const std::string& cell) {
    if (cell == "true" || cell == "TRUE") return true;
    if (cell == "false" || cell == "FALSE") return false;
    throw ParseError{"not a boolean: '" + cell + "'"};
}

void need(const std::vector<std::string>& d, size_t n, const char* what) {
    if (d.size() != n) throw ParseError{std::string(what) + ": expected " + std::to_string(n) + " values"};
}

// one line (action name removed); replay.h:116-463 for the layouts, replay.cpp:5-101 for the
// scaling of locations and radii by float(2^SCENE_DEPTH) / float(2^REPLAY_DEPTH)
hd_replay_action parse_action(const std::string& name, const std::vector<std::string>& d, float scale) {
    hd_replay_action a{};
    auto loc3 = [&](size_t i, float out[3]) {
        for (int k = 0; k < 3; ++k) out[k] = to_f(d[i + k]) * scale;
    };
    if (name == "SetLocation") {
        need(d, 3, "SetLocation");
        a.type = HD_REPLAY_SET_LOCATION;
        for (int k = 0; k < 3; ++k) a.location[k] = (double)to_f(d[k]);
    } else if (name == "SetRotation") {
        need(d, 9, "SetRotation");
        a.type = HD_REPLAY_SET_ROTATION;
        for (int k = 0; k < 9; ++k) a.rotation[k] = (double)to_f(d[k]);
    } else if (name == "SetToolParameters") {
        need(d, 11, "SetToolParameters");
        a.type = HD_REPLAY_SET_TOOL;
        for (int k = 0; k < 9; ++k) (void)to_i(d[k]);
        (void)to_f(d[9]);
        (void)to_i(d[10]);
    } else if (name == "EditSphere" || name == "EditCube") {
        need(d, 5, name.c_str());
        a.type = HD_REPLAY_EDIT;
        const bool add = to_bool(d[4]);
        const bool sphere = name == "EditSphere";
        a.edit.kind = sphere ? (add ? HD_EDIT_SPHERE_ADD : HD_EDIT_SPHERE_CARVE) : (add ? HD_EDIT_BOX_ADD : HD_EDIT_BOX_CARVE);
        loc3(0, a.edit.center);
        a.edit.radius = to_f(d[3]) * scale;
    } else if (name == "EditPaint" || name == "EditFill") {
        need(d, 4, name.c_str());
        a.type = HD_REPLAY_EDIT;
        a.edit.kind = name == "EditPaint" ? HD_EDIT_SPHERE_PAINT : HD_EDIT_FILL;
        loc3(0, a.edit.center);
        a.edit.radius = to_f(d[3]) * scale;
    } else if (name == "EditCopy") {
        if (d.size() != 10 && d.size() != 21) throw ParseError{"EditCopy: expected 10 or 21 values"};
        a.type = HD_REPLAY_EDIT;
        a.edit.kind = HD_EDIT_COPY;
        loc3(0, a.edit.center);
        loc3(3, a.edit.source);
        loc3(6, a.edit.dest);
        a.edit.radius = to_f(d[9]) * scale;
        if (d.size() == 21) {
            for (int k = 0; k < 9; ++k) a.edit.transform[k] = (double)to_f(d[10 + k]);
            a.edit.swirl = to_bool(d[19]) ? 1u : 0u;
            a.edit.swirl_period = to_f(d[20]);
        } else {
            for (int k = 0; k < 9; ++k) a.edit.transform[k] = k % 4 == 0 ? 1.0 : 0.0;  // Matrix3x3()
            a.edit.swirl_period = 1.f;
        }
    } else if (name == "Undo" || name == "Redo" || name == "EndFrame") {
        need(d, 0, name.c_str());
        a.type = name == "Undo" ? HD_REPLAY_UNDO : name == "Redo" ? HD_REPLAY_REDO : HD_REPLAY_END_FRAME;
    } else {
        throw ParseError{"unknown action '" + name + "'"};
    }
    return a;
}

}  // namespace
}  // namespace hd

using namespace hd;

extern "C" {

hd_status hd_replay_parse(const char* text, uint64_t length, uint32_t scene_depth, uint32_t replay_depth,
                          hd_replay** out) {
    if (!out || (!text && length)) return fail(nullptr, HD_ERR_INVALID, "bad arguments");
    if (scene_depth > 31 || replay_depth > 31) return fail(nullptr, HD_ERR_INVALID, "depth above 31");
    *out = nullptr;
    // replay.cpp:5
    const float scale = float(1u << scene_depth) / float(1u << replay_depth);
    auto* r = new hd_replay();
    std::vector<std::string> cells;
    uint64_t line = 1;
    try {
        for (uint64_t i = 0; i < length; ++i) {
            const char c = text[i];
            if (c == '\n') {
                if (cells.empty()) throw ParseError{"empty line"};
                const std::string name = cells.front();
                cells.erase(cells.begin());
                while (!cells.empty() && cells.back().empty()) cells.pop_back();
                hd_replay_action a = parse_action(name, cells, scale);
                a.frame = r->frames;
                if (a.type == HD_REPLAY_END_FRAME) r->frames++;
                r->actions.push_back(a);
                cells.clear();
                ++line;
            } else if (c == ',') {
                cells.emplace_back();
            } else {
                if (cells.empty()) cells.emplace_back();
                cells.back().push_back(c);
            }
        }
        if (!cells.empty()) throw ParseError{"last line does not end with a newline"};
    } catch (const ParseError& e) {
        delete r;
        return fail(nullptr, HD_ERR_INVALID, "replay line " + std::to_string(line) + ": " + e.msg);
    }
    *out = r;
    return HD_OK;
}

hd_status hd_replay_actions(const hd_replay* replay, const hd_replay_action** actions, uint64_t* n_actions,
                            uint32_t* n_frames) {
    if (!replay) return fail(nullptr, HD_ERR_INVALID, "null replay");
    if (actions) *actions = replay->actions.data();
    if (n_actions) *n_actions = replay->actions.size();
    if (n_frames) *n_frames = replay->frames;
    return HD_OK;
}

void hd_replay_free(hd_replay* replay) { delete replay; }

}  // extern "C"
