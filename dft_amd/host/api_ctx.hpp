// api_ctx.hpp -- the per-thread GPU context behind the reference-shaped C++
// API, and host <-> device copies of std::vector<double> fields.
#pragma once

#include <vector>

#include "gpu.hpp"

namespace dftapi {

// Context of the calling thread on the device chosen by dft::setDevice (default 0), created on first use.
dft_ctx* ctx();
// Device copy of a host vector (at least one element allocated).
dftc::Buf up(const std::vector<double>& v);
// The first n doubles of a device buffer.
std::vector<double> down(const dftc::Buf& b, size_t n);

// CartesianGrid (the reference's class or its mirror) -> grid descriptor.
template <class G>
dft_grid to_grid(const G& g) {
    dft_grid d;
    if (dft_grid_init(&d, (int64_t)g.getNx(), (int64_t)g.getNy(), (int64_t)g.getNz(), g.getXMin(), g.getXMax(),
                      g.getYMin(), g.getYMax(), g.getZMin(), g.getZMax()) != DFT_OK)
        throw std::invalid_argument("CartesianGrid requires valid spatial bounds.");
    return d;
}

}  // namespace dftapi
