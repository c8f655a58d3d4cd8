// Lipid templates: the bead count, heads, bonds, angle terms and soup-init
// offsets of one lipid of each architecture, the steps of mol_build
// (cel/engine/oercel.c:479-576) in this library's own code.  Host only.
#pragma once
#include <cmath>
#include <cstring>

#include "../../include/oercel_mi.h"

namespace oc {

constexpr int kMaxMB = 2 * OC_MAXMLEN;  // bonds / angle terms per template

struct MolTmpl {
    int arch, mlen, nH, nT;
    int nb, na;
    int bi[kMaxMB], bj[kMaxMB];
    double br0[kMaxMB];
    int ai[kMaxMB], aj[kMaxMB], ak[kMaxMB], akind[kMaxMB];  // aj: the middle bead
    double aK[kMaxMB], ac0[kMaxMB];
    double px[OC_MAXMLEN], py[OC_MAXMLEN];  // template coordinates (centroid 0)
    double ix[OC_MAXMLEN], iy[OC_MAXMLEN];  // offsets of each bead from bead 0

    void bond(int i, int j, double r0) {
        bi[nb] = i;
        bj[nb] = j;
        br0[nb] = r0;
        nb++;
    }
    // kind 0: U = K (1 - cos); kind 1: U = K (cos - c0)^2; K = 0 terms are left out
    void angle(int i, int j, int k, int kind, double K, double c0) {
        if (K == 0.0) return;
        ai[na] = i;
        aj[na] = j;
        ak[na] = k;
        akind[na] = kind;
        aK[na] = K;
        ac0[na] = c0;
        na++;
    }
    void chain(const int *c, int n, double r0, double ka) {
        for (int b = 0; b + 1 < n; b++) bond(c[b], c[b + 1], r0);
        for (int b = 0; b + 2 < n; b++) angle(c[b], c[b + 1], c[b + 2], 0, ka, 0.0);
    }
};

// Returns false (with *why set) for a template that does not fit OC_MAXMLEN.
inline bool build_template(const oc_params &p, MolTmpl &M, const char **why) {
    memset(&M, 0, sizeof M);
    const double r0 = p.r0;
    const int nT = p.lip_nT;
    M.arch = p.lip_arch;
    M.nT = nT;
    if (nT < 1) {
        *why = "lip_nT must be >= 1";
        return false;
    }
    int ca[OC_MAXMLEN + 2], cb[OC_MAXMLEN + 2];
    switch (p.lip_arch) {
    case OC_ARCH_LINEAR:
        M.nH = p.lip_nH;
        M.mlen = p.lip_nH + nT;
        if (p.lip_nH < 1 || M.mlen > OC_MAXMLEN) {
            *why = "lipid length must be 2..12 with >= 1 head and tail";
            return false;
        }
        for (int b = 0; b + 1 < M.mlen; b++) M.bond(b, b + 1, r0);
        for (int b = 0; b + 2 < M.mlen; b++) M.angle(b, b + 1, b + 2, 0, p.k_angle, 0.0);
        for (int b = 0; b < M.mlen; b++) {
            M.px[b] = (b - 0.5 * (M.mlen - 1)) * r0;
            M.py[b] = 0.0;
        }
        break;
    case OC_ARCH_PI: {
        // heads 0-1 bonded; tail A on head 0: 2..nT+1, tail B on head 1: nT+2..2nT+1
        M.nH = 2;
        M.mlen = 2 + 2 * nT;
        if (M.mlen > OC_MAXMLEN) break;
        const double h = p.r0_head;
        M.bond(0, 1, h);
        ca[0] = 0;
        cb[0] = 1;
        for (int k = 1; k <= nT; k++) {
            ca[k] = 1 + k;
            cb[k] = 1 + nT + k;
        }
        M.chain(ca, nT + 1, r0, p.k_angle);
        M.chain(cb, nT + 1, r0, p.k_angle);
        for (int m = 1; m <= p.lip_tie && m <= nT; m++) M.bond(ca[m], cb[m], h);
        M.angle(1, 0, ca[1], 1, p.k_head, p.cos_head);
        M.angle(0, 1, cb[1], 1, p.k_head, p.cos_head);
        M.py[0] = -0.5 * h;
        M.py[1] = 0.5 * h;
        for (int k = 1; k <= nT; k++) {
            M.px[ca[k]] = k * r0;
            M.py[ca[k]] = -0.5 * h;
            M.px[cb[k]] = k * r0;
            M.py[cb[k]] = 0.5 * h;
        }
        break;
    }
    case OC_ARCH_Y:
        // one head (0) carrying tail A: 1..nT and tail B: nT+1..2nT
        M.nH = 1;
        M.mlen = 1 + 2 * nT;
        if (M.mlen > OC_MAXMLEN) break;
        ca[0] = cb[0] = 0;
        for (int k = 1; k <= nT; k++) {
            ca[k] = k;
            cb[k] = nT + k;
        }
        M.chain(ca, nT + 1, r0, p.k_angle);
        M.chain(cb, nT + 1, r0, p.k_angle);
        for (int m = 1; m <= p.lip_tie && m <= nT; m++) M.bond(ca[m], cb[m], 1.2 * r0);
        M.angle(ca[1], 0, cb[1], 1, p.k_head, p.cos_head);
        for (int k = 1; k <= nT; k++) {
            M.px[ca[k]] = 0.8 * r0 + (k - 1) * r0;
            M.py[ca[k]] = -0.6 * r0;
            M.px[cb[k]] = 0.8 * r0 + (k - 1) * r0;
            M.py[cb[k]] = 0.6 * r0;
        }
        break;
    case OC_ARCH_BOLA: {
        // head pair 0-1, single head 2 on the far end; tails 0 -> 3..nT+2 -> 2
        // and 1 -> nT+3..2nT+2 -> 2
        M.nH = 3;
        M.mlen = 3 + 2 * nT;
        if (M.mlen > OC_MAXMLEN) break;
        const double h = p.r0_head;
        M.bond(0, 1, h);
        ca[0] = 0;
        cb[0] = 1;
        for (int k = 1; k <= nT; k++) {
            ca[k] = 2 + k;
            cb[k] = 2 + nT + k;
        }
        ca[nT + 1] = cb[nT + 1] = 2;
        M.chain(ca, nT + 2, r0, p.k_angle);
        M.chain(cb, nT + 2, r0, p.k_angle);
        for (int m = 1; m <= p.lip_tie && m <= nT; m++) M.bond(ca[m], cb[m], h * (1.0 - (double)m / (nT + 1)));
        M.angle(1, 0, ca[1], 1, p.k_head, p.cos_head);
        M.angle(0, 1, cb[1], 1, p.k_head, p.cos_head);
        const double X = sqrt((nT + 1) * r0 * (nT + 1) * r0 - 0.25 * h * h);
        M.py[0] = -0.5 * h;
        M.py[1] = 0.5 * h;
        M.px[2] = X;
        for (int k = 1; k <= nT; k++) {
            const double f = (double)k / (nT + 1);
            M.px[ca[k]] = f * X;
            M.py[ca[k]] = -0.5 * h * (1.0 - f);
            M.px[cb[k]] = f * X;
            M.py[cb[k]] = 0.5 * h * (1.0 - f);
        }
        break;
    }
    default:
        *why = "lip_arch must be linear, pi, y or bola";
        return false;
    }
    if (M.mlen > OC_MAXMLEN) {
        *why = "lipid template longer than 12 beads";
        return false;
    }
    double mx = 0, my = 0;
    for (int b = 0; b < M.mlen; b++) {
        mx += M.px[b];
        my += M.py[b];
    }
    mx /= M.mlen;
    my /= M.mlen;
    for (int b = 0; b < M.mlen; b++) {
        M.px[b] -= mx;
        M.py[b] -= my;
    }
    for (int b = 0; b < M.mlen; b++) {
        M.ix[b] = M.px[b] - M.px[0];
        M.iy[b] = M.py[b] - M.py[0];
    }
    return true;
}

}  // namespace oc
