#pragma once
// glm-like vector types (only what the module touches)
namespace megamol::stub {
struct vec2 { float x = 0, y = 0; float& operator[](int i) { return (&x)[i]; } float operator[](int i) const { return (&x)[i]; } };
struct vec3 { float x = 0, y = 0, z = 0; float& operator[](int i) { return (&x)[i]; } float operator[](int i) const { return (&x)[i]; } };
struct vec4 { union { struct { float x, y, z, w; }; struct { float r, g, b, a; }; }; vec4() : x(0), y(0), z(0), w(0) {} vec4(float a0, float a1, float a2, float a3) : x(a0), y(a1), z(a2), w(a3) {} };
}  // namespace megamol::stub
