// Element-by-element layout of the level-0 face operator.  See mhs_ebe.h.
#include "mhs_ebe.h"

#include "mhs_kernels.h"

#include <algorithm>
#include <numeric>

namespace mhs {

int greedy_colours(int nc, const int *facepos, const int *nb, std::vector<int> &colour, const int *ijk) {
    if (ijk) {
        colour.assign(nc, 0);
        for (int c = 0; c < nc; ++c) colour[c] = (ijk[3 * (size_t)c] + ijk[3 * (size_t)c + 1] + ijk[3 * (size_t)c + 2]) & 1;
        int ncol = nc > 0 ? 1 : 0;
        for (int c = 0; c < nc; ++c) {
            bool clash = false;
            unsigned long long used = 0;
            for (int h = facepos[c]; h < facepos[c + 1]; ++h) {
                const int o = nb[h];
                if (o < 0) continue;
                if (o < c && colour[o] == colour[c]) clash = true;
                if (colour[o] < 64) used |= 1ull << colour[o];
            }
            if (clash) {
                int k = 2;
                while (k < 64 && (used & (1ull << k))) ++k;
                colour[c] = k;
            }
            ncol = std::max(ncol, colour[c] + 1);
        }
        return ncol;
    }
    colour.assign(nc, -1);
    int ncol = 0;
    for (int c = 0; c < nc; ++c) {
        unsigned long long used = 0;
        for (int h = facepos[c]; h < facepos[c + 1]; ++h) {
            const int o = nb[h];
            if (o >= 0 && colour[o] >= 0 && colour[o] < 64) used |= 1ull << colour[o];
        }
        int k = 0;
        while (k < 64 && (used & (1ull << k))) ++k;
        colour[c] = k;
        ncol = std::max(ncol, k + 1);
    }
    return ncol;
}

std::string ebe_ap_lists(const EbeLayout &E, const Topology &T, const ApPattern &ap, const std::vector<int> &agg,
                         std::vector<int> &ptr, std::vector<int> &src) {
    const long long slots = (long long)ap.col.size();
    if (slots >= 0x7fffffffLL || E.nval >= 0x7fffffffLL) return "too many values for 32-bit lists";
    // slot of (row f, aggregate J): the row's slots are few (about 4)
    auto slot_of = [&](int f, int J) -> long long {
        const int sl = f / kSlice;
        const long long base = ap.slice_ptr[sl] + f % kSlice;
        const int w = int((ap.slice_ptr[sl + 1] - ap.slice_ptr[sl]) / kSlice);
        for (int e = 0; e < w; ++e) {
            const long long q = base + (long long)e * kSlice;
            if (ap.col[q] == J && ap.src_ptr[q + 1] > ap.src_ptr[q]) return q;
        }
        return -1;
    };
    ptr.assign(slots + 1, 0);
    // two passes: counts, then the lists in slot order
    for (int pass = 0; pass < 2; ++pass) {
        std::vector<int> fill;
        if (pass == 1) {
            for (long long q = 0; q < slots; ++q) ptr[q + 1] += ptr[q];
            src.assign(ptr[slots], 0);
            fill.assign(ptr.begin(), ptr.end() - 1);
        }
        for (int c = 0; c < T.nc; ++c) {
            const int s = E.cell_seg[c], l = E.cell_local[c];
            const int f0 = T.facepos[c], n = T.facepos[c + 1] - f0, np = n * (n + 1) / 2;
            for (int i = 0; i < n; ++i) {
                const int f = T.hf_dof[f0 + i];
                if (agg[f] < 0) continue;
                for (int j = 0; j < n; ++j) {
                    const int J = agg[T.hf_dof[f0 + j]];
                    if (J < 0) continue;
                    const long long q = slot_of(f, J);
                    if (q < 0) return "a cell entry outside the A P pattern";
                    if (pass == 0) {
                        ++ptr[q + 1];
                    } else {
                        const int pq = i <= j ? pidx(i, j, n) : pidx(j, i, n);
                        src[fill[q]++] = int(E.seg_val[s] + ebe_slot(l, pq, np));
                    }
                }
            }
        }
    }
    return "";
}

std::string build_ebe(EbeLayout &E, const Topology &T, const int *ijk) {
    E = EbeLayout();
    if (T.has_periodic) return "periodic faces";
    if ((long long)T.nfp > (long long)kEbeRow) return "too many face rows";
    const int nc = T.nc;
    std::vector<int> colour;
    const int ncol = greedy_colours(nc, T.facepos.data(), T.nb.data(), colour, ijk);
    if (ncol > kMaxColours) return "more than 8 colours";
    // instance of every cell
    std::vector<int> inst(nc);
    for (int t = 0; t < T.ninst; ++t)
        for (int c = T.inst_cellpos[t]; c < T.inst_cellpos[t + 1]; ++c) inst[c] = t;
    // segments: cells sorted by (colour, hexahedron or not, instance, face
    // count, index), so that each colour's hexahedra form one block range
    // (launched with the 6-face kernel) followed by the other cells
    auto key = [&](int c) {
        const int n = T.facepos[c + 1] - T.facepos[c];
        return (((long long)colour[c] * 2 + (n == 6 ? 0 : 1)) * T.ninst + inst[c]) * (kMaxFaces + 1) + n;
    };
    std::vector<int> order(nc);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key(a) < key(b); });
    E.ncolours = ncol;
    E.colour_blk.assign(ncol + 1, 0);
    E.cell_seg.assign(nc, -1);
    E.cell_local.assign(nc, -1);
    std::vector<int> seg_colour;
    for (size_t i = 0; i < order.size();) {
        size_t j = i;
        const long long k = key(order[i]);
        while (j < order.size() && key(order[j]) == k) ++j;
        const int c0 = order[i];
        const int n = T.facepos[c0 + 1] - T.facepos[c0];
        const int count = (int)(j - i);
        const int len = (count + kEbeBlock - 1) / kEbeBlock * kEbeBlock;
        const int s = E.nseg++;
        E.seg_n.push_back(n);
        E.seg_inst.push_back(inst[c0]);
        E.seg_count.push_back(count);
        E.seg_len.push_back(len);
        E.seg_blk0.push_back(E.nblk);
        E.seg_val.push_back(E.nval);
        E.seg_dof.push_back(E.ndof);
        seg_colour.push_back(colour[c0]);
        for (int b = 0; b < len / kEbeBlock; ++b) E.blk_seg.push_back(s);
        E.nblk += len / kEbeBlock;
        E.nval += (long long)n * (n + 1) / 2 * len;
        E.ndof += (long long)n * len;
        for (size_t q = i; q < j; ++q) {
            E.cell_seg[order[q]] = s;
            E.cell_local[order[q]] = (int)(q - i);
        }
        i = j;
    }
    for (int k = 0, b = 0; k <= ncol; ++k) {
        while (b < E.nblk && seg_colour[E.blk_seg[b]] < k) ++b;
        E.colour_blk[k] = b;
    }
    E.colour_split.assign(ncol, 0);
    for (int k = 0; k < ncol; ++k) {
        int b = E.colour_blk[k];
        while (b < E.colour_blk[k + 1] && E.seg_n[E.blk_seg[b]] == 6) ++b;
        E.colour_split[k] = b;
    }
    // face rows with the first/last-writer flags, in the 64-cell tiles of the values
    E.dof.assign((size_t)E.ndof, 0u);
    for (int c = 0; c < nc; ++c) {
        const int s = E.cell_seg[c], l = E.cell_local[c];
        const int f0 = T.facepos[c], n = T.facepos[c + 1] - f0;
        for (int j = 0; j < n; ++j) {
            const int h = f0 + j, o = T.nb[h];
            uint32_t d = (uint32_t)T.hf_dof[h];
            if (o < 0 || colour[c] < colour[o]) d |= kEbeFirst;
            if (o < 0 || colour[c] > colour[o]) d |= kEbeLast;
            E.dof[(size_t)E.seg_dof[s] + (size_t)ebe_slot(l, j, n)] = d;
        }
    }
    // partial-sum slots: each instance's blocks (over all colours) contiguous
    std::vector<int> per_inst(T.ninst + 1, 0);
    for (int b = 0; b < E.nblk; ++b) ++per_inst[E.seg_inst[E.blk_seg[b]] + 1];
    E.inst_slot.assign(T.ninst + 1, 0);
    for (int t = 0; t < T.ninst; ++t) E.inst_slot[t + 1] = E.inst_slot[t] + per_inst[t + 1];
    std::vector<int> next(E.inst_slot.begin(), E.inst_slot.end() - 1);
    E.blk_slot.resize(E.nblk);
    for (int b = 0; b < E.nblk; ++b) E.blk_slot[b] = next[E.seg_inst[E.blk_seg[b]]]++;
    E.nslots = E.nblk;
    E.blk_info.resize(4 * (size_t)E.nblk);
    E.blk_base.resize(2 * (size_t)E.nblk);
    for (int b = 0; b < E.nblk; ++b) {
        const int s = E.blk_seg[b], n = E.seg_n[s], local0 = (b - E.seg_blk0[s]) * kEbeBlock;
        E.blk_info[4 * b] = n;
        E.blk_info[4 * b + 1] = E.seg_inst[s];
        E.blk_info[4 * b + 2] = std::min(kEbeBlock, E.seg_count[s] - local0);
        E.blk_info[4 * b + 3] = E.blk_slot[b];
        E.blk_base[2 * b] = E.seg_val[s] + ebe_slot(local0, 0, n * (n + 1) / 2);
        E.blk_base[2 * b + 1] = E.seg_dof[s] + ebe_slot(local0, 0, n);
    }
    return "";
}

}  // namespace mhs
