// Camera and camera-file reader (src/cut/camera.h:39-75, src/cut/camera.cpp:30-75).
#pragma once

#include "../common.h"

namespace cut
{

// Both a perspective and an orthographic camera, as in the reference.
struct Camera
{
    // both
    Mat4f transform;
    Vec3f position;
    // orthographic only
    Vec3f right, up, dir;
    // perspective only
    float scale = 1.f;
    float aspect = 1.f;
    float znear = 0.f; // the reference leaves it uninitialised for orthographic files; here it is 0
    bool orthographic = false;
};

// Text file: "orthographic" flag, 16 floats of the transform (row-major),
// then pos, dir, up, right (orthographic) or scale, znear (perspective; the
// position is column 3 of the transform, aspect = width / height).
Camera load_from_file(string file, int width, int height);

} // namespace cut
