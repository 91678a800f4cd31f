// PNG images of index planes (the reference's PngWriter image outputs,
// src/PngWriter.hpp, src/CurvilinearCSolver_Core.cpp:1438-1752), drawn from a
// structured-grid slice gathered on the device instead of the reference's
// interpolation onto a Cartesian pixel grid.
#pragma once
#include <string>
#include <vector>

namespace mc3d {

enum Colormap { GREYSCALE = 0, RAINBOW = 1, BWR = 2 };

struct ImageOutput {
    int interval = 0;        // every `interval` steps; 0: only at the first step
    int width = 0, height = 0;
    std::string field;       // solver field name
    std::string var_name;    // file stem: <var_name>.<step, 6 digits>.png
    int plane = 2;           // normal direction 0, 1, 2
    double fraction = 0.5;   // plane position along that direction, 0..1
    bool has_range = false;  // fixed colour range, else the plane's min / max
    double vmin = 0.0, vmax = 1.0;
    int colormap = RAINBOW;
};

// 8-bit RGB of f in [0, 1] (src/PngWriter.hpp:270-313 and the greyscale case
// of src/CurvilinearCSolver_Core.cpp:1729-1733).
void colormap_rgb(int cm, double f, int &r, int &g, int &b);

// The image of an (n2, n1) plane (a1 fastest), rows top first: pixel (ip, jp)
// with ip along a1 and jp along a2 (jp = 0 at the bottom) samples the nearest
// index of the pixel centre d/2 + d*ip, d = max((n1-1)/(W+1), (n2-1)/(H+1));
// pixels past the plane take the background colour (73, 175, 205).
std::vector<unsigned char> render_plane(const double *plane, int n1, int n2, const ImageOutput &img);

// 8-bit RGB PNG (zlib-compressed, filter 0).  Returns false when the file
// cannot be written.
bool write_png(const std::string &path, const unsigned char *rgb, int width, int height);

// Plane index of a fraction along a direction of n points.
inline int plane_index(double fraction, int n) {
    int k = (int)(fraction * (n - 1) + 0.5);
    return k < 0 ? 0 : (k > n - 1 ? n - 1 : k);
}

}  // namespace mc3d
