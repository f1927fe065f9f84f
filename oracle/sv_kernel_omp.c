/*
 * CPU restatement, in C with OpenMP, of the statevector kernel's three pair updates
 * (opt_kernel.cpp:71-195: apply_single_gate, apply_cnot, apply_cz) on split real and
 * imaginary arrays, as the FPGA kernel holds the state.  Test infrastructure and CPU
 * baseline only: loaded by tests/ and by bench.py's cpu_baseline leg, never by the
 * product path.  Built by oracle/build_oracle.py (gcc -O3 -fopenmp).
 *
 *   sv_omp_apply_1q(re, im, nq, t, m): for p in [0, 2^(nq-1)): i0 = p with a 0 inserted
 *     at bit t, i1 = i0 | 2^t; (a0, a1) <- M (a0, a1), M the 2x2 complex row-major m[8]
 *     (opt_kernel.cpp:71-117).
 *   sv_omp_apply_cx(re, im, nq, c, t): for p in [0, 2^(nq-2)): zeros inserted at the two
 *     bit positions, control bit set, swap with the target-flipped index (:122-162).
 *   sv_omp_apply_cz(re, im, nq, a, b): both bits set, negate (:167-195).
 *   sv_omp_threads(): OpenMP threads a parallel region uses.
 */
#include <omp.h>
#include <stdint.h>

static inline uint64_t insert_zero(uint64_t p, int b) {
    const uint64_t lo = p & ((1ull << b) - 1);
    return ((p >> b) << (b + 1)) | lo;
}

void sv_omp_apply_1q(double* re, double* im, int nq, int t, const double* m) {
    const int64_t half = (int64_t)1 << (nq - 1);
    const uint64_t st = 1ull << t;
    const double m0r = m[0], m0i = m[1], m1r = m[2], m1i = m[3];
    const double m2r = m[4], m2i = m[5], m3r = m[6], m3i = m[7];
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < half; ++p) {
        const uint64_t i0 = insert_zero((uint64_t)p, t), i1 = i0 | st;
        const double ar = re[i0], ai = im[i0], br = re[i1], bi = im[i1];
        re[i0] = m0r * ar - m0i * ai + m1r * br - m1i * bi;
        im[i0] = m0r * ai + m0i * ar + m1r * bi + m1i * br;
        re[i1] = m2r * ar - m2i * ai + m3r * br - m3i * bi;
        im[i1] = m2r * ai + m2i * ar + m3r * bi + m3i * br;
    }
}

void sv_omp_apply_cx(double* re, double* im, int nq, int c, int t) {
    const int64_t quarter = (int64_t)1 << (nq - 2);
    const int lo = c < t ? c : t, hi = c < t ? t : c;
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < quarter; ++p) {
        const uint64_t base = insert_zero(insert_zero((uint64_t)p, lo), hi);
        const uint64_t i = base | (1ull << c), j = i | (1ull << t);
        const double xr = re[i], xi = im[i];
        re[i] = re[j];
        im[i] = im[j];
        re[j] = xr;
        im[j] = xi;
    }
}

void sv_omp_apply_cz(double* re, double* im, int nq, int a, int b) {
    const int64_t quarter = (int64_t)1 << (nq - 2);
    const int lo = a < b ? a : b, hi = a < b ? b : a;
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < quarter; ++p) {
        const uint64_t i = insert_zero(insert_zero((uint64_t)p, lo), hi) | (1ull << a) | (1ull << b);
        re[i] = -re[i];
        im[i] = -im[i];
    }
}

int sv_omp_threads(void) {
    int n = 1;
#pragma omp parallel
    {
#pragma omp single
        n = omp_get_num_threads();
    }
    return n;
}
