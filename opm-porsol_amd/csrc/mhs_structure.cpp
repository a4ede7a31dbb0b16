// Degree-of-freedom numbering and face-system sparsity (host side, run once per
// initSystemStructure).  See mhs_structure.h.
#include "mhs_structure.h"

#include <algorithm>
#include <cstring>

namespace mhs {

std::string enumerate_instance(int c0, int c1, const int *facepos, const int *nb,
                               int *hf_local, int *nint, int *nf, int *max_ncf) {
    // A face between cells a < b is discovered while visiting a (cells are
    // visited in index order, so "neighbour not yet visited" is b > a,
    // IncompFlowSolverHybrid.hpp:995-1001) and numbered in discovery order.
    // Boundary faces follow in cell / local-face order (:1031-1033).  The
    // half-face on the later cell takes the number of its twin (the
    // reference finds it with std::find in the earlier cell's discovery list,
    // :1036-1056; the first matching entry wins, as here).
    int count = 0, mx = 0;
    for (int c = c0; c < c1; ++c) {
        mx = std::max(mx, facepos[c + 1] - facepos[c]);
        for (int h = facepos[c]; h < facepos[c + 1]; ++h) {
            int n = nb[h];
            if (n < 0) continue;
            if (n < c0 || n >= c1 || n == c) return "neighbour index outside its instance";
            if (n > c) hf_local[h] = count++;
        }
    }
    int total = count;
    for (int c = c0; c < c1; ++c) {
        for (int h = facepos[c]; h < facepos[c + 1]; ++h) {
            int n = nb[h];
            if (n < 0) {
                hf_local[h] = total++;
            } else if (n < c) {
                int found = -1;
                for (int t = facepos[n]; t < facepos[n + 1]; ++t)
                    if (nb[t] == c) { found = t; break; }
                if (found < 0) return "internal face without a twin half-face";
                hf_local[h] = hf_local[found];
            }
        }
    }
    *nint = count;
    *nf = total;
    *max_ncf = mx;
    return "";
}

std::string build_topology(Topology &T, int ninst, const int *inst_cellpos, int nc,
                           const int *facepos, const int *nb, const int *hf_bc_type,
                           const int *hf_partner, const int *hf_family) {
    T = Topology();
    T.nc = nc;
    T.ninst = ninst;
    T.nhf = facepos[nc];
    const int nhf = T.nhf;
    T.facepos.assign(facepos, facepos + nc + 1);
    T.nb.assign(nb, nb + nhf);
    if (inst_cellpos) T.inst_cellpos.assign(inst_cellpos, inst_cellpos + ninst + 1);
    else T.inst_cellpos = {0, nc};
    if (T.inst_cellpos.front() != 0 || T.inst_cellpos.back() != nc) return "inst_cellpos must span [0, nc]";
    for (int c = 0; c < nc; ++c) {
        int n = facepos[c + 1] - facepos[c];
        if (n < 3 || n > kMaxFaces) return "cells must have between 3 and 16 faces";
    }

    // --- numbering per instance -------------------------------------------
    std::vector<int> hf_local(nhf, -1);
    T.inst_nf.resize(ninst);
    T.inst_nint.resize(ninst);
    T.inst_rowbase.resize(ninst + 1);
    T.inst_facebase.resize(ninst + 1);
    T.inst_rowbase[0] = 0;
    T.inst_facebase[0] = 0;
    long long rowbase = 0;
    for (int t = 0; t < ninst; ++t) {
        int c0 = T.inst_cellpos[t], c1 = T.inst_cellpos[t + 1], ni, nf, mx;
        if (c1 < c0) return "inst_cellpos must be non-decreasing";
        std::string e = enumerate_instance(c0, c1, facepos, nb, hf_local.data(), &ni, &nf, &mx);
        if (!e.empty()) return e;
        T.inst_nf[t] = nf;
        T.inst_nint[t] = ni;
        T.max_ncf = std::max(T.max_ncf, mx);
        T.nf_total += nf;
        T.nint_total += ni;
        long long padded = ((long long)std::max(nf, 1) + kRowBlock - 1) / kRowBlock * kRowBlock;
        rowbase += padded;
        if (rowbase > 0x7fffffffLL / 4) return "too many face unknowns";
        T.inst_rowbase[t + 1] = (int)rowbase;
        T.inst_facebase[t + 1] = T.inst_facebase[t] + nf;
    }
    T.nfp = (int)rowbase;
    const int nfp = T.nfp;
    // --- row order: dof -> row within each instance --------------------------
    std::vector<int> row_of_dof(T.nf_total);   // indexed by instance face base + dof
    T.inst_regrow.assign(ninst, 0);
    {
        std::vector<int> fam, len, order;
        for (int t = 0; t < ninst; ++t) {
            const int nf = T.inst_nf[t], fb = T.inst_facebase[t];
            fam.assign(nf, hf_family ? 1 << 30 : 0);
            len.assign(nf, 0);
            for (int c = T.inst_cellpos[t]; c < T.inst_cellpos[t + 1]; ++c)
                for (int h = facepos[c]; h < facepos[c + 1]; ++h) {
                    const int d = hf_local[h];
                    if (hf_family) fam[d] = std::min(fam[d], std::max(0, hf_family[h]));
                    len[d] += facepos[c + 1] - facepos[c];   // row length up to one
                }
            order.resize(nf);
            for (int d = 0; d < nf; ++d) order[d] = d;
            std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return fam[a] < fam[b]; });
            for (int w0 = 0; w0 < nf; w0 += kRowBlock) {
                const int w1 = std::min(nf, w0 + kRowBlock);
                std::stable_sort(order.begin() + w0, order.begin() + w1,
                                 [&](int a, int b) { return len[a] < len[b]; });
            }
            for (int r = 0; r < nf; ++r) row_of_dof[fb + order[r]] = r;
            if (nf > 0) T.inst_regrow[t] = T.inst_rowbase[t] + row_of_dof[fb];
        }
    }
    T.hf_dof.resize(nhf);
    T.hf_face.resize(nhf);
    T.hf_cell.resize(nhf);
    for (int t = 0; t < ninst; ++t)
        for (int c = T.inst_cellpos[t]; c < T.inst_cellpos[t + 1]; ++c)
            for (int h = facepos[c]; h < facepos[c + 1]; ++h) {
                T.hf_face[h] = T.inst_facebase[t] + hf_local[h];
                T.hf_dof[h] = T.inst_rowbase[t] + row_of_dof[T.hf_face[h]];
                T.hf_cell[h] = c;
            }
    T.face_row.resize(T.nf_total);
    T.dof_hf.assign(nfp, -1);
    std::vector<int> dof_hf2(nfp, -1);
    for (int h = 0; h < nhf; ++h) {
        int d = T.hf_dof[h];
        if (T.dof_hf[d] < 0) T.dof_hf[d] = h;
        else dof_hf2[d] = h;
    }
    for (int t = 0; t < ninst; ++t)
        for (int f = 0; f < T.inst_nf[t]; ++f)
            T.face_row[T.inst_facebase[t] + f] = T.inst_rowbase[t] + row_of_dof[T.inst_facebase[t] + f];

    // --- periodic pairing (enumerateBCDof, :1071-1110) ---------------------
    T.hf_twin.assign(nhf, -1);
    T.hf_ppart.assign(nhf, -1);
    T.hf_periodic.assign(nhf, 0);
    T.ptn.assign(nfp, -1);
    for (int h = 0; h < nhf; ++h)
        if (nb[h] >= 0) T.hf_twin[h] = (T.dof_hf[T.hf_dof[h]] == h) ? dof_hf2[T.hf_dof[h]]
                                                                     : T.dof_hf[T.hf_dof[h]];
    if (hf_bc_type && hf_partner) {
        std::vector<int> inst_of_cell(nc);
        for (int t = 0; t < ninst; ++t)
            for (int c = T.inst_cellpos[t]; c < T.inst_cellpos[t + 1]; ++c) inst_of_cell[c] = t;
        for (int h = 0; h < nhf; ++h) {
            if (nb[h] >= 0 || hf_bc_type[h] != 2) continue;
            int p = hf_partner[h];
            if (p < 0 || p >= nhf || nb[p] >= 0) return "periodic half-face without a boundary partner";
            if (p == h) return "periodic half-face paired with itself";
            T.hf_periodic[h] = 1;
            T.hf_ppart[h] = p;
            T.hf_twin[h] = p;
            T.ptn[T.hf_dof[h]] = T.hf_dof[p];
            T.ptn[T.hf_dof[p]] = T.hf_dof[h];
            T.has_periodic = true;
        }
        for (int h = 0; h < nhf; ++h) {
            if (!T.hf_periodic[h]) continue;
            int p = T.hf_ppart[h];
            if (!T.hf_periodic[p] || T.hf_ppart[p] != h) return "periodic partners must be mutual";
            if (inst_of_cell[T.hf_cell[h]] != inst_of_cell[T.hf_cell[p]])
                return "periodic partners in different instances";
        }
    }
    T.hf_col.resize(nhf);
    T.hf_redirect.assign(nhf, 0);
    for (int h = 0; h < nhf; ++h) {
        int d = T.hf_dof[h];
        // the min-dof rule compares reference dof numbers (same instance)
        if (T.hf_periodic[h] && T.hf_face[T.hf_ppart[h]] < T.hf_face[h]) {
            T.hf_redirect[h] = 1;            // addCellContrib :1842-1877 (min-dof rule)
            T.hf_col[h] = T.hf_dof[T.hf_ppart[h]];
        } else {
            T.hf_col[h] = d;
        }
    }

    // --- rows: which half-faces feed which row ------------------------------
    T.row_hf.assign(2 * (size_t)nfp, -1);
    for (int h = 0; h < nhf; ++h) {
        int row = T.hf_col[h];               // row redirection equals column redirection
        int *slot = &T.row_hf[2 * (size_t)row];
        if (slot[0] < 0) slot[0] = h;
        else if (slot[1] < 0) slot[1] = h;
        else return "more than two cell contributions to one face row";
    }
    // keep contributions in half-face order (cell iteration order)
    for (int f = 0; f < nfp; ++f) {
        int *s = &T.row_hf[2 * (size_t)f];
        if (s[0] >= 0 && s[1] >= 0 && s[1] < s[0]) std::swap(s[0], s[1]);
    }
    T.row_cell.assign(2 * (size_t)nfp, -1);
    T.row_loc.assign(2 * (size_t)nfp, 0);
    for (size_t q = 0; q < T.row_hf.size(); ++q) {
        const int h = T.row_hf[q];
        if (h < 0) continue;
        const int c = T.hf_cell[h];
        T.row_cell[q] = c;
        T.row_loc[q] = (uint8_t)(h - facepos[c]);
    }

    // --- SELL-64 pattern --------------------------------------------------
    T.nslices = nfp / kSlice;
    T.row_len.assign(nfp, 1);
    T.dslot.assign(nfp, 0);
    T.pslot.assign(nfp, 255);
    // first pass: per-row column sets (small), kept in a flat scratch
    std::vector<int> cols_flat;
    std::vector<long long> cols_pos(nfp + 1, 0);
    {
        std::vector<int> tmp;
        tmp.reserve(64);
        // count pass
        for (int f = 0; f < nfp; ++f) {
            int n = 1;
            for (int t = 0; t < 2; ++t) {
                int h = T.row_hf[2 * (size_t)f + t];
                if (h < 0) continue;
                int c = T.hf_cell[h];
                n += facepos[c + 1] - facepos[c];
            }
            if (T.ptn[f] >= 0) n += 1;
            cols_pos[f + 1] = cols_pos[f] + n;
        }
        cols_flat.resize(cols_pos[nfp]);
        for (int f = 0; f < nfp; ++f) {
            int *dst = &cols_flat[cols_pos[f]];
            int n = 0;
            dst[n++] = f;
            for (int t = 0; t < 2; ++t) {
                int h = T.row_hf[2 * (size_t)f + t];
                if (h < 0) continue;
                int c = T.hf_cell[h];
                for (int j = facepos[c]; j < facepos[c + 1]; ++j) dst[n++] = T.hf_col[j];
            }
            if (T.ptn[f] >= 0) dst[n++] = T.ptn[f];
            std::sort(dst, dst + n);
            int u = (int)(std::unique(dst, dst + n) - dst);
            if (u > kMaxRowLen) return "face row too long";
            T.row_len[f] = (uint8_t)u;
            for (int e = 0; e < u; ++e) {
                if (dst[e] == f) T.dslot[f] = (uint8_t)e;
                if (T.ptn[f] >= 0 && dst[e] == T.ptn[f]) T.pslot[f] = (uint8_t)e;
            }
        }
    }
    T.slice_ptr.assign(T.nslices + 1, 0);
    for (int s = 0; s < T.nslices; ++s) {
        int w = 0;
        for (int l = 0; l < kSlice; ++l) w = std::max(w, (int)T.row_len[s * kSlice + l]);
        T.slice_ptr[s + 1] = T.slice_ptr[s] + (long long)w * kSlice;
    }
    T.col.assign(T.slice_ptr[T.nslices], 0);
    for (int s = 0; s < T.nslices; ++s) {
        long long base = T.slice_ptr[s];
        int w = (int)((T.slice_ptr[s + 1] - base) / kSlice);
        for (int l = 0; l < kSlice; ++l) {
            int f = s * kSlice + l;
            const int *src = &cols_flat[cols_pos[f]];
            int len = T.row_len[f];
            for (int e = 0; e < w; ++e) T.col[base + (long long)e * kSlice + l] = (e < len) ? src[e] : f;
        }
    }

    // --- per-cell slot maps -------------------------------------------------
    T.sq_pos.assign(nc + 1, 0);
    T.pk_pos.assign(nc + 1, 0);
    for (int c = 0; c < nc; ++c) {
        long long n = facepos[c + 1] - facepos[c];
        T.sq_pos[c + 1] = T.sq_pos[c] + n * n;
        T.pk_pos[c + 1] = T.pk_pos[c] + n * (n + 1) / 2;
    }
    // component-major: entry (r, j) of cell c at (r * max_ncf + j) * nc + c
    const long long mf = T.max_ncf;
    T.slotmap.assign((size_t)(mf * mf) * nc, 255);
    for (int c = 0; c < nc; ++c) {
        int f0 = facepos[c], n = facepos[c + 1] - f0;
        for (int r = 0; r < n; ++r) {
            int row = T.hf_col[f0 + r];
            const int *rc = &cols_flat[cols_pos[row]];
            int len = T.row_len[row];
            for (int j = 0; j < n; ++j) {
                int cj = T.hf_col[f0 + j];
                const int *p = std::lower_bound(rc, rc + len, cj);
                if (p == rc + len || *p != cj) return "slot lookup failed";
                T.slotmap[(size_t)(r * mf + j) * nc + c] = (uint8_t)(p - rc);
            }
        }
    }

    // --- workgroup -> instance map -------------------------------------------
    int nblk = nfp / kRowBlock;
    T.block_seg.resize(nblk);
    T.seg_blk.resize(ninst + 1);
    for (int t = 0; t < ninst; ++t) {
        T.seg_blk[t] = T.inst_rowbase[t] / kRowBlock;
        for (int b = T.inst_rowbase[t] / kRowBlock; b < T.inst_rowbase[t + 1] / kRowBlock; ++b)
            T.block_seg[b] = t;
    }
    T.seg_blk[ninst] = nblk;
    return "";
}

}  // namespace mhs
