// Host-built tables of the step: force-kernel spectrum, FFT twiddles, the
// coincident-pair normals and the liquid kernel constants.
#pragma once

#include <complex>
#include <vector>

namespace gs {

struct GreenKey {
    int mw, mh, pw, ph;
    float aspect, G, sigmaCells, windowFloor;
    bool operator==(const GreenKey& o) const {
        return mw == o.mw && mh == o.mh && pw == o.pw && ph == o.ph && aspect == o.aspect &&
               G == o.G && sigmaCells == o.sigmaCells && windowFloor == o.windowFloor;
    }
};

// Interleaved complex float G(k) for a mw x mh mesh at unit total mass: the
// quarter kx <= mw / 2, ky <= mh / 2, rows of mw / 2 + 1 (the rest follows by
// symmetry: G(-kx, ky) = conj G(kx, ky), G(kx, -ky) = -conj G(kx, ky)).
void green_spectrum(const GreenKey& key, std::vector<float>& out);
// exp(-2 pi i m / length), m < length, interleaved complex float.
void fft_twiddles(int length, std::vector<float>& out);
// Forward FP64 2D DFT (exp(-i ...), unscaled) of a w x h row-major mesh.
void fft2d_f64(std::vector<std::complex<double>>& values, int w, int h);
// cos/sin of the 65536 fallback angles, interleaved.
void fallback_normals(std::vector<float>& out);

// Position-based-fluid kernel constants in pixels (K/LiquidKernel.kt).
struct LiquidConstants {
    float h;       // kernel radius
    float poly6;   // 4 / (pi h^8)
    float spiky;   // -30 / (pi h^5)
    float rest;    // rest density of the hexagonal packing
    float relax;   // constraint relaxation
};
LiquidConstants liquid_constants(float diameter, float radiusDiameters);

}  // namespace gs
