// Host-side planning for the light-field projector: Fourier shapes, radix
// plans and the orbit/item tables of the multiply-accumulate kernel.
//
// Items.  Every half-plane output bin k (k1 <= fX/2) of one plane needs the
// stored spectra at k, (-k0, k1) [row flip], (k1, k0) and (-k1, k0)
// [transposes; square grids only].  Grouping those bins into one item lets
// the kernel read each stored value once and use it for all of them:
//   square     quad  P0=(p0,p1) P1=(-p0,p1) P2=(p1,p0) P3=(-p1,p0),
//              base p in D = {0 <= p1 <= p0 <= f/2}
//   non-square pair  P0=(p0,p1) P1=(-p0,p1), p0 <= fY/2, p1 <= fX/2
// Every half-plane bin is the slot of exactly one item (slots of the same item
// may coincide on the boundaries of D; only the first writes its bin).
//
// Orbits.  The small FFTs S_{b,a} are periodic with the reduced grid r = f/N,
// so the kernel needs S only at the residues (k mod r) of its items' bins.
// Items are grouped by the orbit of their base residue under the symmetry
// group of the item (the dihedral group of the reduced square grid, or the
// sign flips of each axis for non-square grids): one orbit has at most 8
// residues and about N*N items, so a workgroup holds the N*N small-FFT values
// of <= 8 residues in LDS and streams the spectra of its items exactly once.
//
// Aliases.  A bin k = kt + m*r is written with kt the centred residue
// (|kt| < r/2) and m its alias; for the residue r/2 (r even) kt = 0 and m is
// a half integer.  Under this convention negating k negates m exactly, which
// is what lets all flip/transposition phases reduce to per-offset constants
// (folded into the s-table) and per-bin constants (the four classes).
#include "flf_plan.h"

#include <algorithm>
#include <map>

namespace flf {

static int max_prime_factor(int n) {
    int m = 1;
    for (int p = 2; p * p <= n; ++p)
        while (n % p == 0) { m = std::max(m, p); n /= p; }
    if (n > 1) m = std::max(m, n);
    return m;
}

int smooth5_at_least(int n) {
    int m = std::max(n, 1);
    while (m > 1 && max_prime_factor(m) > 5) ++m;
    return m;
}

int fourier_len(int s1, int s2, int N) {
    int need = s1 + s2 - 1;
    int red = (need + N - 1) / N;
    return N * smooth5_at_least(red);
}

bool make_radix_plan(int L, RadixPlan& rp, bool wide, int order) {
    rp.L = L;
    rp.n = 0;
    int m = L;
    // Block-level transforms use radices 2..5; wave-level ones also take the
    // composite radices 8 and 6 first, one pass (an LDS round trip) fewer at
    // most lengths used here.  Radices 9 and 10 were tried: their butterflies
    // raised the register count of every wave-FFT kernel past the occupancy
    // its LDS allows, with spills.  Orders 2 and 3 change the radices of
    // wave-level plans only: the block-level kernels size their butterfly
    // registers from the plan's radices (fft_batch_cap).
    const int narrow[4] = {4, 2, 3, 5};
    const int wider[4][6] = {{8, 6, 4, 5, 3, 2}, {8, 6, 4, 5, 3, 2}, {6, 4, 5, 3, 2, 0}, {5, 3, 8, 6, 4, 2}};
    if (order < 0 || order >= FLF_RADIX_ORDERS) return false;
    const int* list = wide ? wider[order] : narrow;
    const int nlist = wide ? 6 : 4;
    for (int k = 0; k < nlist; ++k) {
        const int R = list[k];
        if (R == 0) break;
        while (m % R == 0) {
            if (rp.n >= FLF_MAX_RADIX_PASSES) return false;
            rp.r[rp.n++] = R;
            m /= R;
        }
    }
    if (order == 1)
        for (int i = 0, j = rp.n - 1; i < j; ++i, --j) std::swap(rp.r[i], rp.r[j]);
    // One odd prime factor above 5 (a lenslet pitch such as 7, 13 or 19) is
    // taken as the last pass, the only position where a radix-P butterfly
    // writes onto its own inputs (see stockham_pass_prime, wave_pass_prime).
    if (m > 5 && m <= FLF_PRIME_RADIX_MAX && max_prime_factor(m) == m) {
        if (rp.n >= FLF_MAX_RADIX_PASSES) return false;
        rp.r[rp.n++] = m;
        m = 1;
    }
    return m == 1;
}

// Half-unit alias of coordinate x on an axis of length N*r.
static int alias2(int x, int r, int N) {
    int k = x % r, m = x / r;
    if (2 * k < r) return (2 * m) % (2 * N);
    if (2 * k > r) return (2 * (m + 1)) % (2 * N);
    return (2 * m + 1) % (2 * N);
}

std::string build_grid_tables(int N, int fY, int fX, GridTables& gt, int items_per_thread) {
    if (N % 2 == 0 || N < 3 || N > FLF_MAX_N) return "Nnum must be odd and in [3, 31]";
    if (fY % N || fX % N) return "Fourier grid must be a multiple of Nnum";
    if (items_per_thread < 1 || items_per_thread > 4) return "items per thread must be 1..4";
    gt = GridTables();
    gt.ipt = items_per_thread;
    gt.fY = fY;
    gt.fX = fX;
    gt.rY = fY / N;
    gt.rX = fX / N;
    gt.rfX = fX / 2 + 1;
    gt.square = (fY == fX);
    gt.slots = gt.square ? 4 : 2;
    const int h = (N + 1) / 2, c = N / 2;

    // Stored spectra order (must match the category loops of k_mac).
    gt.u_bu.push_back(c); gt.u_au.push_back(c);
    if (gt.square) {
        for (int b = 0; b < c; ++b) { gt.u_bu.push_back(b); gt.u_au.push_back(b); }   // (b, b)
        for (int b = 0; b < c; ++b) { gt.u_bu.push_back(b); gt.u_au.push_back(c); }   // (b, c)
        for (int b = 0; b < c; ++b)
            for (int a = b + 1; a < c; ++a) { gt.u_bu.push_back(b); gt.u_au.push_back(a); }
    } else {
        for (int b = 0; b < c; ++b) { gt.u_bu.push_back(b); gt.u_au.push_back(c); }   // (b, c)
        for (int a = 0; a < c; ++a) { gt.u_bu.push_back(c); gt.u_au.push_back(a); }   // (c, a) = transpose of (a, c)
        for (int b = 0; b < c; ++b)
            for (int a = 0; a < c; ++a) { gt.u_bu.push_back(b); gt.u_au.push_back(a); }
    }
    gt.nu = (int)gt.u_bu.size();
    if (gt.nu != (gt.square ? h * (h + 1) / 2 : h * h)) return "internal: spectrum count";

    // Uses of each stored spectrum, in the order of the kernel's use types:
    // identity, both flips, row flip (N-1-b), column flip (N-1-a), then the
    // same four for the transposed PSF (py_light_field.cpp:1514-1560).
    const int NN = N * N;
    gt.nuse = gt.square ? 8 : 4;
    int present = 0;
    for (int u = 0; u < gt.nu; ++u) {
        const int bu = gt.u_bu[u], au = gt.u_au[u], bb = N - 1 - bu, ab = N - 1 - au;
        const bool mx = bu != c, my = au != c, tr = gt.square && au != bu;
        int o[8] = {bu * N + au,
                    (mx && my) ? bb * N + ab : NN,
                    mx ? bb * N + au : NN,
                    my ? bu * N + ab : NN,
                    tr ? au * N + bu : NN,
                    (tr && my && mx) ? ab * N + bb : NN,
                    (tr && my) ? ab * N + bu : NN,
                    (tr && mx) ? au * N + bb : NN};
        for (int q = 0; q < gt.nuse; ++q) {
            gt.uoff.push_back(o[q]);
            present += o[q] < NN;
        }
    }
    if (present != NN) return "internal: uses do not cover every offset once";

    const int rY = gt.rY, rX = gt.rX, rfX = gt.rfX;
    auto canon = [&](int k0, int k1) -> long long {
        long long best = -1;
        int a0[2] = {k0 % rY, (rY - k0 % rY) % rY};
        int a1[2] = {k1 % rX, (rX - k1 % rX) % rX};
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) {
                long long v = (long long)a0[i] * rX + a1[j];
                if (best < 0 || v < best) best = v;
                if (gt.square) {
                    long long w = (long long)a1[j] * rX + a0[i];
                    if (w < best) best = w;
                }
            }
        return best;
    };

    // Collect item bases per orbit (ordered by orbit key for determinism).
    std::vector<std::vector<std::pair<int, int>>> orbit_items((size_t)rY * rX);
    if (gt.square) {
        const int half = fY / 2;
        for (int p0 = 0; p0 <= half; ++p0)
            for (int p1 = 0; p1 <= p0; ++p1)
                orbit_items[canon(p0 % rY, p1 % rX)].push_back({p0, p1});
    } else {
        for (int p0 = 0; p0 <= fY / 2; ++p0)
            for (int p1 = 0; p1 <= fX / 2; ++p1)
                orbit_items[canon(p0 % rY, p1 % rX)].push_back({p0, p1});
    }

    gt.max_res = 0;
    for (size_t key = 0; key < orbit_items.size(); ++key) {
        auto& bases = orbit_items[key];
        if (bases.empty()) continue;
        // Residue list of the orbit in canonical order: the residues of the
        // first item's slots (its base residue k and the row-negated /
        // transposed ones), then their negatives.  The small-FFT kernel stores
        // S at the first half only; the MAC gets the second half by
        // conjugation (S(-k) = conj S(k) for real input).
        std::vector<std::pair<int, int>> res;
        {
            const int k0 = bases[0].first % rY, k1 = bases[0].second % rX;
            const int n0 = (rY - k0) % rY, n1 = (rX - k1) % rX;
            if (gt.square) {
                res = {{k0, k1}, {n0, k1}, {k1, k0}, {n1, k0}, {n0, n1}, {k0, n1}, {n1, n0}, {k1, n0}};
            } else {
                res = {{k0, k1}, {n0, k1}, {n0, n1}, {k0, n1}};
            }
        }
        std::vector<ItemDesc> its;
        its.reserve(bases.size());
        for (auto& bp : bases) {
            int p0 = bp.first, p1 = bp.second;
            int K0[4], K1[4];
            K0[0] = p0; K1[0] = p1;
            K0[1] = (fY - p0) % fY; K1[1] = p1;
            if (gt.square) {
                K0[2] = p1; K1[2] = p0;
                K0[3] = (fY - p1) % fY; K1[3] = p0;
            }
            ItemDesc it{};
            it.writemask = 0;
            for (int j = 0; j < 4; ++j) {
                if (j >= gt.slots) {
                    it.pos[j] = it.pos[0]; it.res[j] = it.res[0]; it.m2[j] = it.m2[0]; it.n2[j] = it.n2[0];
                    continue;
                }
                if (K1[j] < 0 || K1[j] >= rfX) return "internal: slot outside half plane";
                it.pos[j] = K0[j] * rfX + K1[j];
                bool dup = false;
                for (int i = 0; i < j; ++i) dup |= (it.pos[i] == it.pos[j]);
                if (!dup) it.writemask |= (1 << j);
                std::pair<int, int> rr(K0[j] % rY, K1[j] % rX);
                int ri = -1;
                for (size_t q = 0; q < res.size() && ri < 0; ++q)
                    if (res[q] == rr) ri = (int)q;
                if (ri < 0) return "internal: slot residue outside its orbit";
                it.res[j] = (unsigned char)ri;
                it.m2[j] = (unsigned char)alias2(K0[j], rY, N);
                it.n2[j] = (unsigned char)alias2(K1[j], rX, N);
            }
            its.push_back(it);
        }
        gt.max_res = std::max(gt.max_res, (int)res.size());
        gt.real_items += (long long)its.size();
        // Order by residue tuple (residue class): the matrix-core MAC works on
        // tiles of one class, and runs of ipt items share their residues.  For
        // ipt > 1, pad every class to a multiple of ipt with non-writing copies.
        const int ipt = gt.ipt;
        std::stable_sort(its.begin(), its.end(), [](const ItemDesc& x, const ItemDesc& y) {
            for (int j = 0; j < 4; ++j)
                if (x.res[j] != y.res[j]) return x.res[j] < y.res[j];
            return false;
        });
        if (ipt > 1) {
            std::vector<ItemDesc> padded;
            size_t i0 = 0;
            while (i0 < its.size()) {
                size_t i1 = i0;
                while (i1 < its.size() && std::equal(its[i1].res, its[i1].res + 4, its[i0].res)) ++i1;
                for (size_t q = i0; q < i1; ++q) padded.push_back(its[q]);
                size_t n = i1 - i0;
                while (n % ipt) {
                    ItemDesc d = its[i0];
                    d.writemask = 0;
                    padded.push_back(d);
                    ++n;
                }
                i0 = i1;
            }
            its.swap(padded);
        }
        const size_t cmax = (size_t)(FLF_CHUNK_MAX / ipt) * ipt;
        for (size_t s = 0; s < its.size(); s += cmax) {
            ChunkDesc ch{};
            ch.item_base = (long long)gt.items.size();
            ch.n_items = (int)std::min<size_t>(cmax, its.size() - s);
            ch.n_res = (int)res.size();
            for (int q = 0; q < FLF_MAX_RES; ++q) {
                ch.res0[q] = q < ch.n_res ? res[q].first : 0;
                ch.res1[q] = q < ch.n_res ? res[q].second : 0;
            }
            gt.items.insert(gt.items.end(), its.begin() + s, its.begin() + s + ch.n_items);
            // Tiles: runs of one residue class, at most FLF_TILE items each.
            ch.tile_base = (int)gt.tiles.size();
            int i0 = 0;
            while (i0 < ch.n_items) {
                const ItemDesc& first = its[s + i0];
                int i1 = i0;
                while (i1 < ch.n_items && i1 - i0 < FLF_TILE &&
                       std::equal(its[s + i1].res, its[s + i1].res + 4, first.res))
                    ++i1;
                TileDesc td{};
                td.start = i0;
                td.count = i1 - i0;
                for (int j = 0; j < 4; ++j) td.res[j] = first.res[j];
                gt.tiles.push_back(td);
                i0 = i1;
            }
            ch.n_tiles = (int)gt.tiles.size() - ch.tile_base;
            if (ch.n_tiles > FLF_MAX_TILES) return "internal: too many residue-class tiles in a chunk";
            gt.chunks.push_back(ch);
        }
    }
    // Coverage check: every half-plane bin written exactly once.
    std::vector<unsigned char> hit((size_t)fY * rfX, 0);
    for (auto& it : gt.items)
        for (int j = 0; j < gt.slots; ++j)
            if ((it.writemask >> j) & 1) {
                if (hit[it.pos[j]]) return "internal: bin written twice";
                hit[it.pos[j]] = 1;
            }
    for (auto v : hit)
        if (!v) return "internal: bin not covered";
    return "";
}

}  // namespace flf
