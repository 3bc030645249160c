// Host-side scene model: what the .sdf loader leaves behind and the
// compiled flat tape. Mirrors the reference Scene (source/sdf/scene.h:743-900)
// for the parts the render hot path reads.
#pragma once

#include "kodanuki.h"

#include <cstdint>
#include <string>
#include <vector>

namespace kdn {

// A step as the compiler emits it: the reference Step
// (source/sdf/program.h:56-63) with its value in double.
struct Step {
    uint8_t op;
    uint32_t a, b, c;
    double value;
    uint32_t leaf;
};

struct Program {
    std::vector<Step> steps;
    uint32_t result = 0;
};

struct Leaf {
    uint32_t primitive;
    uint32_t material;
    double args[8];
    double linear[3][3];
    double origin[3];
    double stretch;
};

enum Join : uint32_t { JoinLeaf = 0, JoinMerge = 1, JoinCut = 2, JoinClip = 3 };

struct Node {
    Join join;
    double radius;
    uint32_t leaf;
    std::vector<uint32_t> children;
};

struct Scene {
    std::vector<Leaf> leaves;
    std::vector<Node> nodes;
    std::vector<uint32_t> objects;        // root node of each object
    std::vector<kdn_material> materials;
    std::vector<kdn_light> lights;
    kdn_camera camera;
    kdn_render_settings render;
    std::vector<Program> programs;        // one per object after the union split
    std::vector<uint32_t> programMaterial;
};

// Load .sdf text (the subset the render path's scenes use). Returns false
// with "line:column: message" in error.
bool loadScene(const std::string& text, const std::string& baseDir, Scene& scene, std::string& error);

// Compile the object programs: the reference objectPrograms with its
// union split (source/sdf/scene.cpp:274-345).
void compileScene(Scene& scene);

// Remove steps the result does not read (source/sdf/program.cpp:362-407).
Program pruneProgram(const Program& program);

// One leaf alone over world space, as the highlight survey walks it
// (leafProgram, source/sdf/scene.cpp:284-288).
Program leafProgram(const Scene& scene, uint32_t leaf);

int operandCount(uint8_t op);

}  // namespace kdn
