/*
 * FP64 CPU oracle for the real-subspace EPG state machine.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing on the product path links, loads or calls
 * this file; tests/, the smoke check and the bench's CPU-baseline leg use it
 * as the checker.
 *
 * It restates, in double precision and with forward-mode tangents carried as
 * plain arrays, the per-atom recursion of the reference's real-subspace kernels:
 *   - forward:  src/torchsim/_epg_cpu.cpp:3041-3128 (simulate_real_range)
 *   - tangents: src/torchsim/_epg_cpu.cpp:3136-3285 (simulate_real_jvp_range)
 *   - shift:    src/torchsim/_epg_cpu.cpp:3007-3027 (shift_real; top order
 *               dropped, minus[S-1] = 0, plus[0] = -minus[0])
 *   - diffusion weights: src/torchsim/_epg_cpu.cpp:677-715 (l^2 longitudinal,
 *               l^2 + l + 1/3 transverse, b = rate * dt)
 *   - event bits: src/torchsim/_epg_cpu.cpp:36-41
 *
 * Representation (reference comment at _epg_cpu.cpp:3029-3038): with every
 * pulse and sample on one phase axis, F+ = e^{i phi} i a, F- = e^{-i phi} i b,
 * Z = c stay real; the recorded sample is i * m0 * a0. This oracle returns
 * m0 * a0 (the imaginary part), and its tangents.
 *
 * Pinned by: the reference's benchmark checksums (benchmarks/results/
 * torchsim-*.json), the reference's own FP32 CPU kernel built from source,
 * an isochromat (Bloch) brute force at full state count, central finite
 * differences for every tangent, and closed forms (first sample after an
 * inversion, Ernst steady state under spoiling) -- see tests/test_oracle*.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum {
    PRE_SHIFT = 1,
    POST_SHIFT = 2,
    INVERSION = 4,
    SPOIL_AFTER = 8,
    SHIFT_AFTER = 16,
    RECORD = 32,
};

/* Tissue tangent rows, per tangent direction, each (atoms,). */
enum { TAN_T1 = 0, TAN_T2, TAN_M0, TAN_B1, TAN_IE, TAN_DIFF, TAN_TISSUE };
/* Event tangent rows, per tangent direction, each (events,). */
enum { TAN_DURATION = 0, TAN_FLIP, TAN_EVENT };

static void shift_real(double* plus, double* minus, int64_t s) {
    for (int64_t k = 0; k + 1 < s; ++k) minus[k] = minus[k + 1];
    minus[s - 1] = 0.0;
    for (int64_t k = s - 1; k > 0; --k) plus[k] = plus[k - 1];
    plus[0] = -minus[0];
}

static double tissue_at(const double* p, int64_t stride, int64_t atom, double fallback) {
    return p ? p[stride ? atom : 0] : fallback;
}

/*
 * One train for every atom. Tissue pointers may be NULL (feature absent:
 * b1 = m0 = inversion efficiency = 1, diffusion rate = 0); `per_atom` says
 * whether each given buffer has one entry per atom (1) or one in total (0),
 * in the order t1, t2, m0, b1, ie, diffusion.
 *
 * tan_tissue: (tangents, TAN_TISSUE, atoms) or NULL; tan_events: (tangents,
 * TAN_EVENT, events) or NULL. signal: (atoms, outputs). tangent_out:
 * (atoms, tangents, outputs) or NULL.
 */
int epg_oracle_real(
    int64_t atoms, int64_t events, int64_t states, int64_t outputs,
    const double* t1, const double* t2, const double* m0, const double* b1,
    const double* ie, const double* diffusion, const int32_t* per_atom,
    const double* duration, const int32_t* kind, const double* flip,
    const uint8_t* action, const int32_t* output_index,
    int64_t tangents, const double* tan_tissue, const double* tan_events,
    double* signal, double* tangent_out)
{
    if (atoms < 0 || events < 0 || states < 1 || outputs < 0 || tangents < 0) return 1;
    const int64_t S = states, T = tangents;
    double* buf = (double*)calloc((size_t)(3 * S * (1 + T) + 2 * S * (1 + T)), sizeof(double));
    if (!buf) return 2;
    double* a = buf;                 /* (1+T, S): value row then tangent rows */
    double* b = a + S * (1 + T);
    double* c = b + S * (1 + T);
    double* dl = c + S * (1 + T);    /* longitudinal damping (1+T, S) */
    double* dtr = dl + S * (1 + T);  /* transverse damping  (1+T, S) */
    double* dv = (double*)calloc((size_t)(6 * (1 + T)), sizeof(double));
    if (!dv) { free(buf); return 2; }

    for (int64_t atom = 0; atom < atoms; ++atom) {
        memset(a, 0, sizeof(double) * (size_t)(3 * S * (1 + T)));
        c[0] = 1.0;
        const double vt1 = tissue_at(t1, per_atom[0], atom, 1.0);
        const double vt2 = tissue_at(t2, per_atom[1], atom, 1.0);
        const double vm0 = tissue_at(m0, per_atom[2], atom, 1.0);
        const double vb1 = tissue_at(b1, per_atom[3], atom, 1.0);
        const double vie = tissue_at(ie, per_atom[4], atom, 1.0);
        const double vdf = tissue_at(diffusion, per_atom[5], atom, 0.0);
        /* dv[k*6 + q]: tangent k of tissue quantity q. */
        for (int64_t k = 0; k < T; ++k)
            for (int q = 0; q < TAN_TISSUE; ++q)
                dv[k * 6 + q] = tan_tissue ? tan_tissue[(k * TAN_TISSUE + q) * atoms + atom] : 0.0;

        for (int64_t e = 0; e < events; ++e) {
            const double dt = duration[e];
            /* Relaxation over the interval that precedes the event. */
            const double r1 = 1000.0 / vt1, r2 = 1000.0 / vt2;
            const double e1 = exp(-r1 * dt), e2 = exp(-r2 * dt);
            const double bf = vdf * dt;
            for (int64_t s = 0; s < S; ++s) {
                const double l = (double)s;
                dl[s] = exp(-(l * l) * bf);
                dtr[s] = exp(-(l * l + l + 1.0 / 3.0) * bf);
            }
            /* tangents of e1, e2 and the damping, then apply */
            for (int64_t k = 0; k < T; ++k) {
                const double ddt = tan_events ? tan_events[(k * TAN_EVENT + TAN_DURATION) * events + e] : 0.0;
                const double dt1 = dv[k * 6 + TAN_T1], dt2 = dv[k * 6 + TAN_T2];
                const double ddf = dv[k * 6 + TAN_DIFF];
                /* d(r1*dt) = -1000 dt1/t1^2 * dt + r1 * ddt */
                const double de1 = -e1 * (-1000.0 * dt1 / (vt1 * vt1) * dt + r1 * ddt);
                const double de2 = -e2 * (-1000.0 * dt2 / (vt2 * vt2) * dt + r2 * ddt);
                const double dbf = ddf * dt + vdf * ddt;
                double* ak = a + (1 + k) * S;
                double* bk = b + (1 + k) * S;
                double* ck = c + (1 + k) * S;
                for (int64_t s = 0; s < S; ++s) {
                    const double l = (double)s;
                    const double ddl = -dl[s] * (l * l) * dbf;
                    const double ddtr = -dtr[s] * (l * l + l + 1.0 / 3.0) * dbf;
                    const double tr = e2 * dtr[s], dtrv = de2 * dtr[s] + e2 * ddtr;
                    const double lo = e1 * dl[s], dlo = de1 * dl[s] + e1 * ddl;
                    ak[s] = ak[s] * tr + a[s] * dtrv;
                    bk[s] = bk[s] * tr + b[s] * dtrv;
                    ck[s] = ck[s] * lo + c[s] * dlo;
                }
                ck[0] -= de1;
            }
            for (int64_t s = 0; s < S; ++s) {
                const double tr = e2 * dtr[s], lo = e1 * dl[s];
                a[s] *= tr;
                b[s] *= tr;
                c[s] *= lo;
            }
            c[0] += 1.0 - e1;

            const unsigned act = action[e];
            if (act & PRE_SHIFT)
                for (int64_t k = 0; k <= T; ++k) shift_real(a + k * S, b + k * S, S);
            if (kind[e] == 1) {
                if (act & INVERSION) {
                    const double eff = -vie;
                    for (int64_t k = 0; k < T; ++k) {
                        const double deff = -dv[k * 6 + TAN_IE];
                        double* ck = c + (1 + k) * S;
                        for (int64_t s = 0; s < S; ++s) ck[s] = ck[s] * eff + c[s] * deff;
                    }
                    for (int64_t s = 0; s < S; ++s) c[s] *= eff;
                } else {
                    const double alpha = flip[e] * vb1;
                    const double co = cos(alpha), si = sin(alpha);
                    const double ch = 0.5 * (1.0 + co), sh = 0.5 * (1.0 - co), hs = 0.5 * si;
                    for (int64_t k = 0; k < T; ++k) {
                        const double dfl = tan_events ? tan_events[(k * TAN_EVENT + TAN_FLIP) * events + e] : 0.0;
                        const double dal = dfl * vb1 + flip[e] * dv[k * 6 + TAN_B1];
                        const double dco = -si * dal, dsi = co * dal;
                        const double dch = -0.5 * si * dal, dsh = 0.5 * si * dal, dhs = 0.5 * co * dal;
                        double* ak = a + (1 + k) * S;
                        double* bk = b + (1 + k) * S;
                        double* ck = c + (1 + k) * S;
                        for (int64_t s = 0; s < S; ++s) {
                            const double p = a[s], m = b[s], z = c[s];
                            const double dp = ak[s], dm = bk[s], dz = ck[s];
                            ak[s] = ch * dp + dch * p + sh * dm + dsh * m - si * dz - dsi * z;
                            bk[s] = sh * dp + dsh * p + ch * dm + dch * m + si * dz + dsi * z;
                            ck[s] = hs * dp + dhs * p - hs * dm - dhs * m + co * dz + dco * z;
                        }
                    }
                    for (int64_t s = 0; s < S; ++s) {
                        const double p = a[s], m = b[s], z = c[s];
                        a[s] = ch * p + sh * m - si * z;
                        b[s] = sh * p + ch * m + si * z;
                        c[s] = hs * p - hs * m + co * z;
                    }
                }
            } else if (kind[e] == 2 && (act & RECORD)) {
                const int64_t o = output_index[e];
                if (o >= 0 && o < outputs) {
                    signal[atom * outputs + o] = vm0 * a[0];
                    if (tangent_out)
                        for (int64_t k = 0; k < T; ++k)
                            tangent_out[(atom * T + k) * outputs + o] =
                                dv[k * 6 + TAN_M0] * a[0] + vm0 * a[(1 + k) * S];
                }
            }
            if (act & POST_SHIFT)
                for (int64_t k = 0; k <= T; ++k) shift_real(a + k * S, b + k * S, S);
            if (act & SPOIL_AFTER) {
                for (int64_t k = 0; k <= T; ++k) {
                    memset(a + k * S, 0, sizeof(double) * (size_t)S);
                    memset(b + k * S, 0, sizeof(double) * (size_t)S);
                }
            } else if (act & SHIFT_AFTER) {
                for (int64_t k = 0; k <= T; ++k) shift_real(a + k * S, b + k * S, S);
            }
        }
    }
    free(dv);
    free(buf);
    return 0;
}
