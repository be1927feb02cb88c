// Seeded decaying-isotropic-turbulence initial condition.
//
// The reference ships no turbulence generator (it restarts from a file,
// src/main.f:69-71, src/out.f:598-692), so this is new: a random-phase,
// solenoidal velocity with shell spectrum E(k) ~ k^4 exp(-2 (k/k0)^2),
// modes with |k| >= N/2 (in grid units) removed, scaled so that the
// per-component rms sqrt(<u.u>/3) equals urms; rho = 1 and
// p = 1/(gamma Mt^2).  Random numbers come from a counter-based generator
// keyed on (seed, mode index), so the field is bit-identical per seed
// regardless of threading.
#include <cmath>
#include <complex>
#include <cstdint>
#include <vector>

#include "st_internal.h"

namespace st {

namespace {

typedef std::complex<double> cplx;

uint64_t splitmix(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

double uni(uint64_t key)
{
    return ((splitmix(key) >> 11) + 0.5) * (1.0 / 9007199254740992.0);   // (0,1)
}

double gauss(uint64_t key)
{
    const double u1 = uni(key * 2 + 0), u2 = uni(key * 2 + 1);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}

int smallest_factor(int n)
{
    for (int p = 2; p * p <= n; p++)
        if (n % p == 0) return p;
    return n;
}

// Recursive mixed-radix inverse DFT (sign +1) of x (length n, stride s) into y.
void idft(const cplx *x, int n, int s, cplx *y, cplx *work)
{
    if (n == 1) { y[0] = x[0]; return; }
    const int p = smallest_factor(n);
    const int m = n / p;
    if (p == n) {                       // prime length: direct sum
        for (int k = 0; k < n; k++) {
            cplx acc = 0;
            for (int j = 0; j < n; j++) acc += x[(long)j * s] * std::polar(1.0, 2.0 * M_PI * (double)((long)j * k % n) / n);
            y[k] = acc;
        }
        return;
    }
    // y_r = IDFT of x[r + p*j], j = 0..m-1, for residues r = 0..p-1
    for (int r = 0; r < p; r++) idft(x + (long)r * s, m, s * p, work + (long)r * m, y);
    for (int k = 0; k < m; k++) {
        for (int q = 0; q < p; q++) {
            const int kk = k + q * m;
            cplx acc = 0;
            for (int r = 0; r < p; r++) acc += work[(long)r * m + k] * std::polar(1.0, 2.0 * M_PI * (double)((long)r * kk % n) / n);
            y[kk] = acc;
        }
    }
}

// In-place inverse DFT along one axis of a 3-D array a[i0 + n0*(i1 + n1*i2)].
void idft_axis(std::vector<cplx> &a, const int n[3], int axis)
{
    const long len = n[axis];
    const long stride = axis == 0 ? 1 : (axis == 1 ? n[0] : (long)n[0] * n[1]);
    const long total = (long)n[0] * n[1] * n[2];
    const long lines = total / len;
#pragma omp parallel
    {
        std::vector<cplx> in(len), out(len), work(len * 2);
#pragma omp for schedule(static)
        for (long l = 0; l < lines; l++) {
            long base;
            if (axis == 0) base = l * len;
            else if (axis == 1) base = (l % n[0]) + (l / n[0]) * (long)n[0] * n[1];
            else base = l;
            for (long i = 0; i < len; i++) in[i] = a[base + i * stride];
            idft(in.data(), (int)len, 1, out.data(), work.data());
            for (long i = 0; i < len; i++) a[base + i * stride] = out[i];
        }
    }
}

}  // namespace

// Q(My,Mz,Mx,5) in Fortran order: index j + My*(k + Mz*i) per field.
int ic_isotropic(int nx, int ny, int nz, const double period[3], double k0, double urms,
                 double mt, double gamma, unsigned long long seed, double *Q)
{
    if (nx < 4 || ny < 4 || nz < 4 || k0 <= 0 || urms <= 0 || mt <= 0) return -1;
    // spectral array in the physical layout order (y fastest, then z, then x)
    const int n[3] = {ny, nz, nx};
    const long N = (long)nx * ny * nz;
    std::vector<cplx> U[3] = {std::vector<cplx>(N), std::vector<cplx>(N), std::vector<cplx>(N)};
    const int nmin = std::min(nx, std::min(ny, nz));
    const double kcut = 0.5 * nmin;     // |k| >= N/2 removed (grid units for a 2*pi box)
#pragma omp parallel for schedule(static)
    for (long idx = 0; idx < N; idx++) {
        const int j = (int)(idx % ny), k = (int)((idx / ny) % nz), i = (int)(idx / ((long)ny * nz));
        const int mx = i <= nx / 2 ? i : i - nx;
        const int my = j <= ny / 2 ? j : j - ny;
        const int mz = k <= nz / 2 ? k : k - nz;
        for (int c = 0; c < 3; c++) U[c][idx] = 0;
        if (2 * std::abs(mx) >= nx || 2 * std::abs(my) >= ny || 2 * std::abs(mz) >= nz) continue;
        const double kx = 2.0 * M_PI / period[0] * mx;
        const double ky = 2.0 * M_PI / period[1] * my;
        const double kz = 2.0 * M_PI / period[2] * mz;
        const double kk = std::sqrt(kx * kx + ky * ky + kz * kz);
        if (kk == 0.0) continue;
        // radius in grid units for the cut-off (isotropic cube: same as kk for L = 2 pi)
        const double kg = std::sqrt((double)mx * mx + (double)my * my + (double)mz * mz);
        if (kg >= kcut) continue;
        const double amp = kk * std::exp(-(kk / k0) * (kk / k0));   // sqrt(E(k)/(4 pi k^2)) up to a constant
        const uint64_t key = splitmix(seed) ^ (uint64_t)idx * 0x9E3779B97F4A7C15ULL;
        cplx a[3];
        for (int c = 0; c < 3; c++) a[c] = cplx(gauss(key * 8 + 2 * c), gauss(key * 8 + 2 * c + 1));
        const double kv[3] = {kx, ky, kz};
        const cplx dot = (a[0] * kv[0] + a[1] * kv[1] + a[2] * kv[2]) / (kk * kk);
        for (int c = 0; c < 3; c++) U[c][idx] = (a[c] - dot * kv[c]) * amp;
    }
    for (int c = 0; c < 3; c++)
        for (int ax = 0; ax < 3; ax++) idft_axis(U[c], n, ax);
    double s2 = 0.0;
    for (int c = 0; c < 3; c++)
        for (long l = 0; l < N; l++) s2 += U[c][l].real() * U[c][l].real();
    const double rms = std::sqrt(s2 / (3.0 * N));
    if (!(rms > 0)) return -2;
    const double sc = urms / rms;
    const double p = 1.0 / (gamma * mt * mt);
#pragma omp parallel for schedule(static)
    for (long l = 0; l < N; l++) {
        const double u = U[0][l].real() * sc, v = U[1][l].real() * sc, w = U[2][l].real() * sc;
        Q[l] = u;
        Q[N + l] = v;
        Q[2 * N + l] = w;
        Q[3 * N + l] = 1.0;
        Q[4 * N + l] = p / (gamma - 1.0) + 0.5 * (u * u + v * v + w * w);
    }
    return 0;
}

// Planar shock with upstream turbulence for the non-periodic x grid
// (x_i = period_x i/(nx-1)).  Upstream of x_s = frac L the gas is at rest
// state rho1 = 1, p1 = 1/gamma (sound speed 1) moving at u1 = M; downstream
// the Rankine-Hugoniot state of a normal shock of Mach M:
//   rho2/rho1 = (gamma+1) M^2 / ((gamma-1) M^2 + 2),
//   p2/p1 = 1 + 2 gamma/(gamma+1) (M^2 - 1),  u2 = u1 rho1/rho2,
// joined by phi = (1 + tanh((x - x_s)/width))/2.  The isotropic field of
// ic_isotropic (k0, u_rms) is added to the velocity upstream, scaled by
// (1 - phi).  states[0..4] and [5..9] return the conserved upstream and
// downstream states (rho u, rho v, rho w, rho, rho E).
int ic_shock(int nx, int ny, int nz, const double period[3], double mach, double frac, double width,
             double k0, double urms, double gamma, unsigned long long seed, double *Q, double *states)
{
    if (nx < 16 || mach <= 1.0 || frac <= 0.0 || frac >= 1.0 || width <= 0.0) return -1;
    const long N = (long)nx * ny * nz;
    std::vector<double> T(urms > 0.0 ? 5 * N : 0);
    if (urms > 0.0) {
        const int rc = ic_isotropic(nx, ny, nz, period, k0, urms, 1.0, gamma, seed, T.data());
        if (rc) return rc;
    }
    const double M2 = mach * mach;
    const double r1 = 1.0, p1 = 1.0 / gamma, u1 = mach;
    const double r2 = r1 * (gamma + 1.0) * M2 / ((gamma - 1.0) * M2 + 2.0);
    const double p2 = p1 * (1.0 + 2.0 * gamma / (gamma + 1.0) * (M2 - 1.0));
    const double u2 = u1 * r1 / r2;
    const double L = period[0], xs = frac * L;
    if (states) {
        const double e1 = p1 / (gamma - 1.0) + 0.5 * r1 * u1 * u1, e2 = p2 / (gamma - 1.0) + 0.5 * r2 * u2 * u2;
        const double s[10] = {r1 * u1, 0.0, 0.0, r1, e1, r2 * u2, 0.0, 0.0, r2, e2};
        for (int i = 0; i < 10; i++) states[i] = s[i];
    }
#pragma omp parallel for schedule(static)
    for (long l = 0; l < N; l++) {
        const int i = (int)(l / ((long)ny * nz));
        const double x = L * (double)i / (double)(nx - 1);
        const double phi = 0.5 * (1.0 + std::tanh((x - xs) / width));
        const double rho = r1 + (r2 - r1) * phi, p = p1 + (p2 - p1) * phi;
        double u = u1 + (u2 - u1) * phi, v = 0.0, w = 0.0;
        if (urms > 0.0) {
            u += (1.0 - phi) * T[l];
            v += (1.0 - phi) * T[N + l];
            w += (1.0 - phi) * T[2 * N + l];
        }
        Q[l] = rho * u;
        Q[N + l] = rho * v;
        Q[2 * N + l] = rho * w;
        Q[3 * N + l] = rho;
        Q[4 * N + l] = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w);
    }
    return 0;
}

}  // namespace st
