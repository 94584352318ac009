// Spectrum of the P3M long-range force kernel, built once on the host in
// FP64 (S/kernel_init.comp:16-46, S/convolve.comp:24-47,
// K/GravityPipeline.kt:1341-1357).
//
// The reference evaluates the Gaussian-split kernel Kx, Ky in FP32 and
// transforms them on the GPU whenever the total mass changes. Here both are
// evaluated and transformed in FP64 at unit mass, the CIC deconvolution and
// the 1/(W H) of the unscaled inverse are folded in, and the two spectra are
// packed as G = s (Kx^ + i Ky^): the spectral multiply then needs one complex
// product, and 1/M is applied at run time, so a change of body count never
// rebuilds the table.
#include <cmath>
#include <complex>
#include <vector>

#include "green.hpp"

namespace gs {

using cd = std::complex<double>;

static void fft_inplace(cd* a, int n, int stride) {
    // Iterative radix-2 decimation in time, forward (exp(-i ...)).
    std::vector<cd> line(n);
    for (int i = 0; i < n; i++) line[i] = a[(size_t)i * stride];
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(line[i], line[j]);
    }
    for (int len = 2; len <= n; len <<= 1) {
        double ang = -2.0 * M_PI / len;
        for (int i = 0; i < n; i += len) {
            for (int k = 0; k < len / 2; k++) {
                cd w(std::cos(ang * k), std::sin(ang * k));
                cd u = line[i + k], v = line[i + k + len / 2] * w;
                line[i + k] = u + v;
                line[i + k + len / 2] = u - v;
            }
        }
    }
    for (int i = 0; i < n; i++) a[(size_t)i * stride] = line[i];
}

static void fft2d(std::vector<cd>& m, int w, int h) {
    for (int y = 0; y < h; y++) fft_inplace(m.data() + (size_t)y * w, w, 1);
    for (int x = 0; x < w; x++) fft_inplace(m.data() + x, h, w);
}

void fft2d_f64(std::vector<std::complex<double>>& values, int w, int h) { fft2d(values, w, h); }

static double sinc(double v) { return std::fabs(v) < 1e-12 ? 1.0 : std::sin(v) / v; }

void green_spectrum(const GreenKey& k, std::vector<float>& out) {
    const int mw = k.mw, mh = k.mh;
    std::vector<cd> kx((size_t)mw * mh), ky((size_t)mw * mh);
    const double aspect = (double)k.aspect;
    const double sigma = (double)k.sigmaCells / (double)k.ph;
    for (int ry = 0; ry < mh; ry++) {
        for (int rx = 0; rx < mw; rx++) {
            int ox = rx <= mw / 2 ? rx : rx - mw;
            int oy = ry <= mh / 2 ? ry : ry - mh;
            double dx = (double)ox * aspect / (double)k.pw;
            double dy = (double)oy / (double)k.ph;
            double r2 = dx * dx + dy * dy;
            size_t id = (size_t)ry * mw + rx;
            if (r2 == 0.0) continue;
            double lw = 1.0 - std::exp(-r2 / (2.0 * sigma * sigma));
            double dirx = rx * 2 == mw ? 0.0 : dx;
            double diry = ry * 2 == mh ? 0.0 : dy;
            double scale = -(double)k.G * lw / r2;  // unit total mass
            kx[id] = cd(scale * dirx / aspect, 0.0);
            ky[id] = cd(scale * diry, 0.0);
        }
    }
    fft2d(kx, mw, mh);
    fft2d(ky, mw, mh);
    // Kx is odd in x and even in y, Ky the reverse (the sampling above keeps
    // that exactly, the Nyquist column and row included), so Kx^ = i a and
    // Ky^ = i b with a odd in kx and even in ky, b the reverse, and
    // G = s (i a - b): G(-kx, ky) = conj G(kx, ky), G(kx, -ky) = -conj G(kx, ky).
    // The quarter kx <= mw / 2, ky <= mh / 2 holds the whole table.
    const int qw = mw / 2 + 1, qh = mh / 2 + 1;
    out.resize((size_t)2 * qw * qh);
    const double inv = 1.0 / ((double)mw * (double)mh);
    for (int ry = 0; ry < qh; ry++) {
        double sy = sinc(M_PI * ry / mh);
        for (int rx = 0; rx < qw; rx++) {
            double sx = sinc(M_PI * rx / mw);
            double window = sx * sx * sy * sy;
            double deconv = 1.0 / std::fmax(window * window, (double)k.windowFloor);
            double s = deconv * inv;
            size_t id = (size_t)ry * mw + rx;
            cd g = s * (kx[id] + cd(0.0, 1.0) * ky[id]);
            size_t q = (size_t)ry * qw + rx;
            out[2 * q] = (float)g.real();
            out[2 * q + 1] = (float)g.imag();
        }
    }
}

void fft_twiddles(int length, std::vector<float>& out) {
    out.resize((size_t)2 * length);
    for (int m = 0; m < length; m++) {
        double a = -2.0 * M_PI * (double)m / (double)length;
        out[2 * m] = (float)std::cos(a);
        out[2 * m + 1] = (float)std::sin(a);
    }
}

void fallback_normals(std::vector<float>& out) {
    // S/collision_cache_solve.comp:105-112: the FP32 angle of each of the
    // 65536 hashes, its cos/sin rounded once from FP64.
    out.resize(2 * 65536);
    const float step = (float)(6.28318530718 / 65536.0);
    for (int h = 0; h < 65536; h++) {
        float angle = (float)h * step;
        out[2 * h] = (float)std::cos((double)angle);
        out[2 * h + 1] = (float)std::sin((double)angle);
    }
}

LiquidConstants liquid_constants(float diameter, float radiusDiameters) {
    // K/LiquidKernel.kt:14-60: 2D poly6 and spiky coefficients at the kernel
    // radius, the rest density as the poly6 sum over the hexagonal packing of
    // bodies one diameter apart (the body itself included), and the
    // relaxation as 1% of that packing's squared spiky-gradient sum.
    LiquidConstants k{};
    const float h = diameter * radiusDiameters;
    const double hd = (double)h;
    const double h5 = hd * hd * hd * hd * hd;
    k.h = h;
    k.poly6 = (float)(4.0 / (M_PI * (h5 * hd * hd * hd)));
    k.spiky = (float)(-30.0 / (M_PI * h5));
    const double d = (double)diameter;
    const double rowStep = d * std::sqrt(3.0) * 0.5;
    const int rows = (int)std::ceil(hd / rowStep) + 1;
    const int columns = (int)std::ceil(hd / d) + 1;
    auto each = [&](auto&& f) {
        for (int r = -rows; r <= rows; r++) {
            double shift = (r & 1) == 0 ? 0.0 : 0.5 * d;
            for (int c = -columns; c <= columns; c++) f(c * d + shift, r * rowStep);
        }
    };
    double rest = 0.0;
    each([&](double x, double y) {
        double q = hd * (double)h - (x * x + y * y);
        rest += q > 0.0 ? (double)k.poly6 * q * q * q : 0.0;
    });
    k.rest = (float)rest;
    double relax = 0.0;
    each([&](double x, double y) {
        double r = std::sqrt(x * x + y * y);
        double g = 0.0;
        if (r > 0.0 && r < hd) {
            double q = hd - r;
            g = -(double)k.spiky * q * q;
        }
        g /= (double)k.rest;
        relax += g * g;
    });
    k.relax = (float)(0.01 * relax);
    return k;
}

}  // namespace gs
