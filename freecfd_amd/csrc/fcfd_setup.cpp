// Host set-up of the residual path: connectivity, geometry, boundary ghosts,
// BC tagging, length scales, interpolation weights and gradient maps, in the
// reference's numbering, followed by a reverse Cuthill-McKee renumbering for
// the device layout.
//
// The connectivity is built with hash/sort/bitmask data structures instead of
// the reference's linear searches, but it reproduces the reference's orderings
// exactly (checked bit-for-bit against the CPU restatement in tests):
//   - local node ids: first touch in cell order (grid_create_elements.cc:35-59)
//   - faces: creation order = (owner cell, local face slot) of the first
//     occurrence, neighbour = the other cell (grid_create_elements.cc:225-384)
//   - cell face lists ascending (first-free-slot filling, :364-377)
//   - neighborCells: node-sharing cells in (cell node, node cell) order incl.
//     self, then boundary ghosts (:146-164, :745-769)
// Geometry uses the reference formulas and operation order (grid.cc:123-178);
// this file is compiled with -ffp-contract=off so it rounds like the reference.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <numeric>
#include <unordered_map>

#include "fcfd_internal.h"

namespace fcfd {
namespace {

struct V3 { double x, y, z; };
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double mag(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
inline V3 unit(V3 a) { return a / mag(a); }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, -a.x * b.z + a.z * b.x, a.x * b.y - a.y * b.x}; }
inline V3 get3(const std::vector<double>& v, int i) { return {v[3 * i], v[3 * i + 1], v[3 * i + 2]}; }
inline void put3(std::vector<double>& v, int i, V3 a) { v[3 * i] = a.x; v[3 * i + 1] = a.y; v[3 * i + 2] = a.z; }

// local face templates, CGNS element node order (grid_create_elements.cc:187-214)
const int kHexF[6][4] = {{0, 3, 2, 1}, {4, 5, 6, 7}, {1, 2, 6, 5}, {0, 4, 7, 3}, {1, 5, 4, 0}, {2, 3, 7, 6}};
const int kPriF[5][4] = {{0, 2, 1, 0}, {3, 4, 5, 0}, {0, 3, 5, 2}, {1, 2, 5, 4}, {0, 1, 4, 3}};
const int kPyrF[5][4] = {{0, 3, 2, 1}, {0, 1, 4, 0}, {1, 2, 4, 0}, {3, 4, 2, 0}, {0, 4, 3, 0}};
const int kTetF[4][3] = {{0, 2, 1}, {1, 2, 3}, {0, 3, 2}, {0, 1, 3}};

int faces_of(int nn) { return nn == 4 ? 4 : nn == 8 ? 6 : 5; }
int face_nodes(int nn, int cf, const int* cn, int* out) {
    switch (nn) {
    case 4: for (int q = 0; q < 3; ++q) out[q] = cn[kTetF[cf][q]]; return 3;
    case 5: { int k = cf < 1 ? 4 : 3; for (int q = 0; q < k; ++q) out[q] = cn[kPyrF[cf][q]]; return k; }
    case 6: { int k = cf < 2 ? 3 : 4; for (int q = 0; q < k; ++q) out[q] = cn[kPriF[cf][q]]; return k; }
    default: for (int q = 0; q < 4; ++q) out[q] = cn[kHexF[cf][q]]; return 4;
    }
}

struct FaceKey {
    int n[4];
    bool operator==(const FaceKey& o) const { return n[0] == o.n[0] && n[1] == o.n[1] && n[2] == o.n[2] && n[3] == o.n[3]; }
};
struct FaceKeyHash {
    size_t operator()(const FaceKey& k) const {
        uint64_t h = 1469598103934665603ull;
        for (int q = 0; q < 4; ++q) { h ^= (uint64_t)(uint32_t)k.n[q]; h *= 1099511628211ull; }
        return (size_t)(h ^ (h >> 29));
    }
};

int lowest_bit(uint32_t m) { return m ? __builtin_ctz(m) : -1; }

// Weighted tri-linear interpolation weights of a point from a stencil of
// cell centroids (Interpolate, src/interpolate/interpolate.cc:36-351).  One
// object serves one pass over faces or nodes: like the reference's members
// weights3/weights4 it keeps the last small-system solution, which a
// singular system (solve() leaves x alone) then reuses.
class Wtli {
public:
    Wtli(int dim, int max_stencil, double skew_tol) : dim_(dim), maxs_(max_stencil), tol_(skew_tol) {}

    // Reorders ids/pts as the reference's sort_stencil(is_internal = true)
    // and truncates them; fills w.  Returns the number of weights: the
    // stencil size for tetra or triangle fits, 2 for a line, 1 for a point.
    int fit(V3 x, std::vector<int>& ids, std::vector<V3>& pts, std::vector<double>& w) {
        keep_two_then_nearest(x, ids, pts);
        const int m = (int)pts.size();
        w.assign(m, 0.0);
        if (tetra(x, pts, w) || tri(x, pts, w)) return m;
        if (m >= 2) {   // interpolate_line (:320-343)
            V3 e1 = pts[1] - pts[0];
            V3 b1 = unit(e1);
            V3 q = dot(x - pts[0], b1) * b1;
            w[1] = dot(q, b1) / mag(e1);
            w[0] = 1. - w[1];
            return 2;
        }
        w[0] = 1.;      // interpolate_point (:345-351)
        return 1;
    }

private:
    int dim_, maxs_;
    double tol_;
    std::array<double, 4> x4_{};
    std::array<double, 3> x3_{};

    // the first two stay; the rest in ascending distance, an equal distance
    // placed before the earlier entry (sort_stencil, :65-125)
    void keep_two_then_nearest(V3 x, std::vector<int>& ids, std::vector<V3>& pts) const {
        const int n = (int)pts.size(), m = std::min(n, maxs_);
        std::vector<int> order(m, 0);
        std::vector<double> best(m, 1.e20);
        if (m > 1) order[1] = 1;
        for (int d = 2; d < n; ++d) {
            const double dd = mag(pts[d] - x);
            for (int q = 2; q < m; ++q) {
                if (dd <= best[q]) {
                    for (int i = m - 1; i > q; --i) { order[i] = order[i - 1]; best[i] = best[i - 1]; }
                    order[q] = d;
                    best[q] = dd;
                    break;
                }
            }
        }
        std::vector<int> ni(m);
        std::vector<V3> np(m);
        for (int i = 0; i < m; ++i) { ni[i] = ids[order[i]]; np[i] = pts[order[i]]; }
        ids.swap(ni);
        pts.swap(np);
    }

    // Gaussian elimination with a row swap only for a pivot below 1e-12
    // (gelimd, src/utilities/utilities.cc:71-101); x unchanged when singular
    template <int K>
    static bool solve(double (&a)[K][K], double (&b)[K], std::array<double, K>& x) {
        constexpr double kEps = 1e-12;
        for (int i = 0; i < K; ++i) {
            double pvt = a[i][i];
            if (std::fabs(pvt) < kEps) {
                int j = i + 1;
                for (; j < K; ++j) { pvt = a[j][i]; if (std::fabs(pvt) >= kEps) break; }
                if (std::fabs(pvt) < kEps) return false;
                for (int c = 0; c < K; ++c) std::swap(a[i][c], a[j][c]);
                std::swap(b[i], b[j]);
            }
            for (int k = i + 1; k < K; ++k) {
                const double t = a[k][i] / pvt;
                for (int c = i + 1; c < K; ++c) a[k][c] -= t * a[i][c];
                b[k] -= t * b[i];
            }
        }
        for (int i = K - 1; i >= 0; --i) {
            double v = b[i];
            for (int c = K - 1; c > i; --c) v -= a[i][c] * x[c];
            x[i] = v / a[i][i];
        }
        return true;
    }

    static void normalise(std::vector<double>& w) {
        double s = 0.;
        for (double v : w) s += v;
        for (double& v : w) v /= s;
    }

    // blended barycentric fits over every admissible tetrahedron (:127-243)
    bool tetra(V3 x, const std::vector<V3>& P, std::vector<double>& w) {
        const int m = (int)P.size();
        if (dim_ != 3 || m < 4) return false;
        bool ok = false;
        for (int s1 = 0; s1 < m - 3; ++s1)
            for (int s2 = s1 + 1; s2 < m - 2; ++s2)
                for (int s3 = s2 + 1; s3 < m - 1; ++s3)
                    for (int s4 = s3 + 1; s4 < m; ++s4) {
                        const V3 e1 = P[s2] - P[s1], e2 = P[s3] - P[s1], e3 = P[s3] - P[s2];
                        const V3 e4 = P[s4] - P[s1], e5 = P[s4] - P[s2], e6 = P[s4] - P[s3];
                        const double edge = (mag(e1) + mag(e2) + mag(e3) + mag(e4) + mag(e5) + mag(e6)) / 6.;
                        V3 nrm = cross(e1, e2);
                        const double area = 0.5 * mag(nrm);
                        nrm = unit(nrm);
                        const double vol = area * std::fabs(dot(e4, nrm)) / 3.;
                        const double regular = std::sqrt(2.) / 12. * edge * edge * edge;
                        if (1. - vol / regular > tol_) continue;   // too skewed
                        const V3 cen = (((P[s1] + P[s2]) + P[s3]) + P[s4]) / 4.;
                        double dist = mag(x - cen) / edge;
                        if (dist < 1.e-8) dist = 1.e-8;
                        const double blend = 1. / (dist * dist);
                        const int s[4] = {s1, s2, s3, s4};
                        double a[4][4], b[4] = {x.x, x.y, x.z, 1.};
                        for (int q = 0; q < 4; ++q) { a[0][q] = P[s[q]].x; a[1][q] = P[s[q]].y; a[2][q] = P[s[q]].z; a[3][q] = 1.; }
                        solve<4>(a, b, x4_);
                        const double sum = ((x4_[0] + x4_[1]) + x4_[2]) + x4_[3];
                        if (std::fabs(sum - 1.) < 1.e-8) {
                            for (int q = 0; q < 4; ++q) w[s[q]] += blend * x4_[q];
                            ok = true;
                        }
                    }
        normalise(w);
        return ok;
    }

    // blended fits over every admissible triangle, in-plane coordinates with
    // the point projected through the origin (:245-318)
    bool tri(V3 x, const std::vector<V3>& P, std::vector<double>& w) {
        const int m = (int)P.size();
        if (dim_ < 2 || m < 3) return false;
        std::fill(w.begin(), w.end(), 0.0);
        bool ok = false;
        for (int s1 = 0; s1 < m - 2; ++s1)
            for (int s2 = s1 + 1; s2 < m - 1; ++s2)
                for (int s3 = s2 + 1; s3 < m; ++s3) {
                    const V3 e1 = P[s2] - P[s1], e2 = P[s3] - P[s1], e3 = P[s3] - P[s2];
                    V3 nrm = cross(e1, e2);
                    const double area = 0.5 * mag(nrm);
                    nrm = unit(nrm);
                    const double edge = (mag(e1) + mag(e2) + mag(e3)) / 3.;
                    if (1. - area / (0.433 * edge * edge) > tol_) continue;   // too skewed
                    V3 q = x - dot(x, nrm) * nrm;
                    const V3 cen = ((P[s1] + P[s2]) + P[s3]) / 3.;
                    double dist = mag(q - cen) / edge;
                    if (dist < 1.e-2) dist = 1.e-2;
                    const double blend = 1. / (dist * dist);
                    const V3 b1 = unit(e1);
                    const V3 b2 = (-1.) * unit(cross(b1, nrm));
                    q = q - P[s1];
                    double a[3][3] = {{0., dot(e1, b1), dot(e2, b1)}, {0., dot(e1, b2), dot(e2, b2)}, {1., 1., 1.}};
                    double b[3] = {dot(q, b1), dot(q, b2), 1.};
                    solve<3>(a, b, x3_);
                    const double sum = (x3_[0] + x3_[1]) + x3_[2];
                    if (std::fabs(sum - 1.) < 1.e-8) {
                        w[s1] += blend * x3_[0]; w[s2] += blend * x3_[1]; w[s3] += blend * x3_[2];
                        ok = true;
                    }
                }
        normalise(w);
        return ok;
    }
};

// std::map<int, double>::insert of (ids[s], w[s]), s < n, appended to a CSR
// row in ascending keys (a repeated key keeps its first weight)
void append_map_row(const std::vector<int>& ids, const std::vector<double>& w, int n, std::vector<int>& idx, std::vector<double>& wt) {
    std::vector<std::pair<int, double>> row;
    for (int s = 0; s < n; ++s) {
        bool dup = false;
        for (auto& e : row) if (e.first == ids[s]) { dup = true; break; }
        if (!dup) row.push_back({ids[s], w[s]});
    }
    std::stable_sort(row.begin(), row.end(), [](const std::pair<int, double>& a, const std::pair<int, double>& b) { return a.first < b.first; });
    for (auto& e : row) { idx.push_back(e.first); wt.push_back(e.second); }
}

}  // namespace

int build_host_grid(const fcfd_mesh* m, const fcfd_options* o, HostGrid& g, std::string& err,
                    bool wall_distance_on_host) {
    if (!m || !o || m->n_cells <= 0 || m->n_nodes <= 0 || m->n_bc < 0 || m->n_bc > FCFD_MAX_BC || o->n_bc != m->n_bc) {
        err = "invalid mesh or options (cell/node counts, n_bc)";
        return FCFD_EINVAL;
    }
    const int N = m->n_cells;
    g.dim = m->dimension;
    g.N = N;
    // ---- local nodes by first touch, cell node lists -------------------------
    std::vector<int> g2l(m->n_nodes, -1);
    g.cell_nptr.assign(N + 1, 0);
    g.cell_nodes.clear();
    g.cell_nodes.reserve(m->cell_ptr[N]);
    g.node.clear();
    int nn_local = 0;
    for (int c = 0; c < N; ++c) {
        int nn = m->cell_ptr[c + 1] - m->cell_ptr[c];
        if (nn != 4 && nn != 5 && nn != 6 && nn != 8) { err = "unsupported cell type"; return FCFD_EMESH; }
        for (int q = 0; q < nn; ++q) {
            int ng = m->cell_nodes[m->cell_ptr[c] + q];
            if (ng < 0 || ng >= m->n_nodes) { err = "node id out of range"; return FCFD_EINVAL; }
            if (g2l[ng] < 0) {
                g2l[ng] = nn_local++;
                g.node.push_back(m->xyz[3 * ng]); g.node.push_back(m->xyz[3 * ng + 1]); g.node.push_back(m->xyz[3 * ng + 2]);
            }
            g.cell_nodes.push_back(g2l[ng]);
        }
        g.cell_nptr[c + 1] = (int)g.cell_nodes.size();
    }
    // cells with repeated nodes (collapsed edges): their faces keep each node
    // once, and a hexahedron with two repeated nodes becomes a prism after the
    // faces are built (grid_create_elements.cc:274-294, 426-467)
    std::vector<char> repeated(N, 0);
    for (int c = 0; c < N; ++c) {
        const int* cn = &g.cell_nodes[g.cell_nptr[c]];
        const int nn = g.cell_nptr[c + 1] - g.cell_nptr[c];
        for (int a = 0; a < nn && !repeated[c]; ++a)
            for (int b = 0; b < a; ++b)
                if (cn[a] == cn[b]) { repeated[c] = 1; break; }
    }
    g.nnode = nn_local;
    // BC membership bitmask per local node
    std::vector<uint32_t> bcmask(nn_local, 0u);
    for (int b = 0; b < m->n_bc; ++b)
        for (int k = m->bc_ptr[b]; k < m->bc_ptr[b + 1]; ++k) {
            int ng = m->bc_nodes[k];
            if (ng >= 0 && ng < m->n_nodes && g2l[ng] >= 0) bcmask[g2l[ng]] |= (1u << b);
        }
    // ---- node -> cells CSR (ascending cells) ----------------------------------
    // (a repeated node lists its cell once, grid_create_elements.cc:126-142)
    auto first_in_cell = [&](int c, int k) {
        for (int j = g.cell_nptr[c]; j < k; ++j) if (g.cell_nodes[j] == g.cell_nodes[k]) return false;
        return true;
    };
    std::vector<int> ncptr(nn_local + 1, 0), ncell;
    for (int c = 0; c < N; ++c)
        for (int k = g.cell_nptr[c]; k < g.cell_nptr[c + 1]; ++k)
            if (!repeated[c] || first_in_cell(c, k)) ncptr[g.cell_nodes[k] + 1]++;
    for (int n = 0; n < nn_local; ++n) ncptr[n + 1] += ncptr[n];
    ncell.resize(ncptr[nn_local]);
    {
        std::vector<int> fill(ncptr.begin(), ncptr.end() - 1);
        for (int c = 0; c < N; ++c)
            for (int k = g.cell_nptr[c]; k < g.cell_nptr[c + 1]; ++k)
                if (!repeated[c] || first_in_cell(c, k)) ncell[fill[g.cell_nodes[k]]++] = c;
    }
    // ---- faces by hashing sorted node tuples --------------------------------
    std::unordered_map<FaceKey, int, FaceKeyHash> fmap;
    fmap.reserve((size_t)N * 4);
    g.fparent.clear(); g.fneigh.clear(); g.fbc.clear(); g.fnptr.assign(1, 0); g.fnodes.clear();
    std::vector<std::vector<int>> cfl(N);
    for (int c = 0; c < N; ++c) {
        int nn = g.cell_nptr[c + 1] - g.cell_nptr[c];
        const int* cn = &g.cell_nodes[g.cell_nptr[c]];
        for (int cf = 0; cf < faces_of(nn); ++cf) {
            int fnod[4];
            int k = face_nodes(nn, cf, cn, fnod);
            if (repeated[c]) {   // keep each node once; a face left with two nodes is not a face
                int u = 0;
                for (int q = 0; q < k; ++q) {
                    bool seen = false;
                    for (int j = 0; j < q; ++j) if (fnod[j] == fnod[q]) { seen = true; break; }
                    if (!seen) fnod[u++] = fnod[q];
                }
                k = u;
                if (k == 2) continue;
            }
            FaceKey key{{-1, -1, -1, -1}};
            for (int q = 0; q < k; ++q) key.n[q] = fnod[q];
            std::sort(key.n, key.n + k);
            auto it = fmap.find(key);
            if (it == fmap.end()) {
                int f = (int)g.fparent.size();
                fmap.emplace(key, f);
                g.fparent.push_back(c);
                g.fneigh.push_back(-1);
                g.fbc.push_back(-1);
                for (int q = 0; q < k; ++q) g.fnodes.push_back(fnod[q]);
                g.fnptr.push_back((int)g.fnodes.size());
                cfl[c].push_back(f);
            } else {
                int f = it->second;
                if (g.fneigh[f] >= 0) { err = "non-manifold face (shared by more than two cells)"; return FCFD_EMESH; }
                g.fneigh[f] = c;
                cfl[c].push_back(f);
            }
        }
    }
    const int F = (int)g.fparent.size();
    g.F = F;
    g.cfptr.assign(N + 1, 0);
    g.cfaces.clear();
    for (int c = 0; c < N; ++c) {
        std::sort(cfl[c].begin(), cfl[c].end());
        for (int f : cfl[c]) g.cfaces.push_back(f);
        g.cfptr[c + 1] = (int)g.cfaces.size();
        std::vector<int>().swap(cfl[c]);
    }
    // ---- BC tag of boundary faces (node-set match with the multi-BC tie-break)
    for (int f = 0; f < F; ++f) {
        if (g.fneigh[f] >= 0) continue;
        uint32_t fm = ~0u;
        for (int k = g.fnptr[f]; k < g.fnptr[f + 1]; ++k) fm &= bcmask[g.fnodes[k]];
        if (m->n_bc < 32) fm &= ((1u << m->n_bc) - 1u);
        int bc = -2;
        if (fm) {
            if (__builtin_popcount(fm) > 1) {
                int c = g.fparent[f];
                uint32_t cm = ~0u;
                for (int k = g.cell_nptr[c]; k < g.cell_nptr[c + 1]; ++k) cm &= bcmask[g.cell_nodes[k]];
                // the reference records the first region containing every cell node,
                // scanning regions in order (grid_create_elements.cc:338-349)
                int cmb = lowest_bit(cm & ((m->n_bc < 32) ? ((1u << m->n_bc) - 1u) : ~0u));
                for (uint32_t r = fm; r; r &= r - 1) {
                    int b = __builtin_ctz(r);
                    if (b != cmb) { bc = b; break; }
                }
            } else bc = lowest_bit(fm);
        }
        if (bc == -2) { err = "boundary condition could not be found for a face"; return FCFD_EBC; }
        g.fbc[f] = bc;
    }
    // ---- hexahedra with two repeated nodes become prisms (:426-467): the node
    // list is read off the cell's triangular faces, in its face order, each
    // rotated to start at a repeated node.  The node-sharing neighbours of
    // interior cells keep the original lists (they are built before the faces
    // in the reference, create_nodes_cells).
    std::vector<int> orig_nptr, orig_nodes;
    {
        bool any = false;
        for (int c = 0; c < N; ++c) any |= repeated[c] != 0;
        if (any) {
            orig_nptr = g.cell_nptr;
            orig_nodes = g.cell_nodes;
            std::vector<int> np(N + 1, 0), nodes;
            nodes.reserve(g.cell_nodes.size());
            for (int c = 0; c < N; ++c) {
                const int* cn = &orig_nodes[orig_nptr[c]];
                const int nn = orig_nptr[c + 1] - orig_nptr[c];
                std::vector<int> rep;
                if (repeated[c])
                    for (int a = 0; a < nn; ++a)
                        for (int b = 0; b < a; ++b)
                            if (cn[a] == cn[b]) rep.push_back(cn[a]);
                if (nn == 8 && rep.size() == 2) {
                    for (int k = g.cfptr[c]; k < g.cfptr[c + 1]; ++k) {
                        const int f = g.cfaces[k];
                        if (g.fnptr[f + 1] - g.fnptr[f] != 3) continue;
                        const int* fn3 = &g.fnodes[g.fnptr[f]];
                        int q = 0;
                        for (; q < 3; ++q) if (fn3[q] == rep[0] || fn3[q] == rep[1]) break;
                        if (q == 3) continue;
                        for (int j = 0; j < 3; ++j) nodes.push_back(fn3[(q + j) % 3]);
                    }
                } else {
                    nodes.insert(nodes.end(), cn, cn + nn);
                }
                np[c + 1] = (int)nodes.size();
            }
            g.cell_nptr.swap(np);
            g.cell_nodes.swap(nodes);
        }
    }
    // ---- areas, normals, centroids, volumes (grid.cc:123-178) ----------------
    g.fc.assign(3 * F, 0.0); g.fn.assign(3 * F, 0.0); g.ft1.assign(3 * F, 0.0); g.farea.assign(F, 0.0);
    auto nd = [&](int n) { return get3(g.node, n); };
    for (int f = 0; f < F; ++f) {
        const int* fnod = &g.fnodes[g.fnptr[f]];
        int k = g.fnptr[f + 1] - g.fnptr[f];
        if (k == 4) {
            V3 nrm = cross(nd(fnod[2]) - nd(fnod[0]), nd(fnod[3]) - nd(fnod[1]));
            g.farea[f] = 0.5 * mag(nrm);
            put3(g.fn, f, unit(nrm));
        } else {
            V3 av{0.0, 0.0, 0.0};
            V3 c0 = nd(fnod[0]);
            for (int q = 1; q < k; ++q) {
                int nx = (q + 1 == k) ? 0 : q + 1;
                av = av + 0.5 * cross(nd(fnod[q]) - c0, nd(fnod[nx]) - c0);
            }
            g.farea[f] = mag(av);
            put3(g.fn, f, unit(av));
        }
        V3 s{0.0, 0.0, 0.0};
        for (int q = 0; q < k; ++q) s = s + nd(fnod[q]);
        V3 fcen = s / (double)k;
        put3(g.fc, f, fcen);
        // face tangent (ns_assemble_linear_system.cc:337-341)
        V3 t1 = (k == 4) ? unit((nd(fnod[0]) + nd(fnod[1])) - (nd(fnod[2]) + nd(fnod[3])))
                         : unit(0.5 * (nd(fnod[0]) + nd(fnod[1])) - fcen);
        put3(g.ft1, f, t1);
    }
    int B = 0;
    for (int f = 0; f < F; ++f) if (g.fbc[f] >= 0) ++B;
    g.B = B;
    const int T = N + B;
    g.cc.assign(3 * T, 0.0); g.vol.assign(T, 0.0);
    double total = 0.0;
    for (int c = 0; c < N; ++c) {
        V3 s{0.0, 0.0, 0.0};
        int nn = g.cell_nptr[c + 1] - g.cell_nptr[c];
        for (int k = g.cell_nptr[c]; k < g.cell_nptr[c + 1]; ++k) s = s + nd(g.cell_nodes[k]);
        V3 cen = s / (double)nn;
        put3(g.cc, c, cen);
        double v = 0.0;
        for (int k = g.cfptr[c]; k < g.cfptr[c + 1]; ++k) {
            int f = g.cfaces[k];
            V3 base = g.farea[f] * get3(g.fn, f);
            if (g.fparent[f] == c) base = {base.x * -1., base.y * -1., base.z * -1.};
            V3 bn = unit(base);
            V3 h = dot(cen - get3(g.fc, f), bn) * bn;
            v += dot(base, h) / 3.;
        }
        g.vol[c] = v;
        total += v;
    }
    g.total_volume = total;
    // ---- boundary ghosts: per BC region, in face order (grid_create_elements.cc:702-743)
    g.cbc.assign(T, -1);
    g.ghost_face.assign(B, -1);
    {
        std::vector<int> cnt(m->n_bc + 1, 0);
        for (int f = 0; f < F; ++f) if (g.fbc[f] >= 0) cnt[g.fbc[f] + 1]++;
        for (int b = 0; b < m->n_bc; ++b) cnt[b + 1] += cnt[b];
        for (int f = 0; f < F; ++f) {
            int b = g.fbc[f];
            if (b < 0) continue;
            int gidx = N + cnt[b]++;
            int p = g.fparent[f];
            V3 pc = get3(g.cc, p), fcen = get3(g.fc, f), nrm = get3(g.fn, f);
            put3(g.cc, gidx, pc + (2. * dot(fcen - pc, nrm)) * nrm);
            g.vol[gidx] = g.vol[p];
            g.cbc[gidx] = b;
            g.ghost_face[gidx - N] = f;
            g.fneigh[f] = gidx;
        }
    }
    // ---- neighborCells (node sharing incl. self), then ghosts ---------------
    {
        std::vector<int> ngptr(nn_local + 1, 0), nghost;
        for (int q = 0; q < B; ++q) {
            int f = g.ghost_face[q];
            for (int k = g.fnptr[f]; k < g.fnptr[f + 1]; ++k) ngptr[g.fnodes[k] + 1]++;
        }
        for (int n = 0; n < nn_local; ++n) ngptr[n + 1] += ngptr[n];
        nghost.resize(ngptr[nn_local]);
        std::vector<int> fill(ngptr.begin(), ngptr.end() - 1);
        for (int q = 0; q < B; ++q) {
            int f = g.ghost_face[q];
            for (int k = g.fnptr[f]; k < g.fnptr[f + 1]; ++k) nghost[fill[g.fnodes[k]]++] = N + q;
        }
        std::vector<std::vector<int>> lists(T);
        std::vector<int> stamp(T, -1);
        const std::vector<int>& onp = orig_nptr.empty() ? g.cell_nptr : orig_nptr;
        const std::vector<int>& onode = orig_nptr.empty() ? g.cell_nodes : orig_nodes;
        for (int c = 0; c < N; ++c) {
            auto& L = lists[c];
            for (int k = onp[c]; k < onp[c + 1]; ++k) {
                int n = onode[k];
                for (int j = ncptr[n]; j < ncptr[n + 1]; ++j) {
                    int c2 = ncell[j];
                    if (stamp[c2] != c) { stamp[c2] = c; L.push_back(c2); }
                }
            }
            for (int k = g.cell_nptr[c]; k < g.cell_nptr[c + 1]; ++k) {
                int n = g.cell_nodes[k];
                for (int j = ngptr[n]; j < ngptr[n + 1]; ++j) {
                    int gg = nghost[j];
                    if (stamp[gg] != c) { stamp[gg] = c; L.push_back(gg); lists[gg].push_back(c); }
                }
            }
        }
        g.nbptr.assign(T + 1, 0);
        g.nb.clear();
        for (int c = 0; c < T; ++c) {
            g.nb.insert(g.nb.end(), lists[c].begin(), lists[c].end());
            g.nbptr[c + 1] = (int)g.nb.size();
        }
    }
    // ---- set_bcs: symmetry flags, face2bc, region areas, ghost centroid fix --
    g.fsym.assign(F, 0); g.face2bc.assign(F, 0); g.bc_area.assign(m->n_bc, 0.0);
    {
        std::vector<int> counter(m->n_bc, 0);
        for (int f = 0; f < F; ++f) {
            int b = g.fbc[f];
            if (b < 0) continue;
            g.face2bc[f] = counter[b]++;
            if (o->bc[b].type == FCFD_BC_SYMMETRY) g.fsym[f] = 1;
        }
        for (int b = 0; b < m->n_bc; ++b) {   // same summation order as set_bcs.cc:86-95
            double a = 0.0;
            for (int f = 0; f < F; ++f) if (g.fbc[f] == b) a += g.farea[f];
            g.bc_area[b] = a;
        }
        for (int f = 0; f < F; ++f) {   // set_bcs.cc:153-159
            int b = g.fbc[f];
            if (b >= 0 && (o->bc[b].type == FCFD_BC_INLET || o->bc[b].type == FCFD_BC_OUTLET)) {
                int p = g.fparent[f];
                put3(g.cc, g.fneigh[f], get3(g.cc, p) + 2. * (get3(g.fc, f) - get3(g.cc, p)));
            }
        }
    }
    // ---- wall distance (set_bcs.cc:178-249) --------------------------------
    g.wdist.assign(T, 1.e20);
    g.fwd.assign(F, 1.e20);
    g.fdiss.assign(F, 0.0);
    if (wall_distance_on_host) {
        std::vector<int> wf;
        for (int f = 0; f < F; ++f) {
            int b = g.fbc[f];
            if (b >= 0 && o->bc[b].type == FCFD_BC_WALL && o->bc[b].kind != FCFD_KIND_SLIP) wf.push_back(f);
        }
        for (int c = 0; c < T; ++c) {
            double d = 1.e20;
            V3 x = get3(g.cc, c);
            for (int f : wf) { double t = mag(x - get3(g.fc, f)); if (d > t) d = t; }
            g.wdist[c] = d;
        }
        for (int f = 0; f < F; ++f) {
            int b = g.fbc[f];
            if (b < 0 || o->bc[b].type != FCFD_BC_WALL || o->bc[b].kind == FCFD_KIND_SLIP) {
                double d = 1.e20;
                V3 fN{0.0, 0.0, 0.0}, x = get3(g.fc, f);
                for (int w : wf) { double t = mag(x - get3(g.fc, w)); if (d > t) { d = t; fN = get3(g.fn, w); } }
                g.fwd[f] = d;
                V3 n = get3(g.fn, f);
                g.fdiss[f] = 1. - std::fabs(fN.x * n.x + fN.y * n.y + fN.z * n.z);
            } else { g.fwd[f] = 0.; g.fdiss[f] = 0.; }
        }
    }
    // ---- length scales (main.cc:456-491) ----------------------------------
    g.lscale.assign(T, 0.0);
    for (int c = 0; c < N; ++c) {
        double ls = 1.e20;
        V3 cen = get3(g.cc, c);
        for (int k = g.cfptr[c]; k < g.cfptr[c + 1]; ++k) {
            int f = g.cfaces[k];
            double h = std::fabs(dot(get3(g.fn, f), get3(g.fc, f) - cen));
            bool skip = g.fbc[f] >= 0 && o->bc[g.fbc[f]].type == FCFD_BC_SYMMETRY;
            if (!skip) ls = (h < ls) ? h : ls;
        }
        g.lscale[c] = ls;
    }
    if (g.dim == 3) g.glscale = std::pow(g.total_volume, 1. / 3.);
    else {
        g.glscale = 0.0;
        for (int b = 0; b < m->n_bc; ++b)
            if (o->bc[b].type == FCFD_BC_SYMMETRY) { double s = std::sqrt(0.5 * g.bc_area[b]); if (g.glscale < s) g.glscale = s; }
    }
    // ---- face interpolation weights (face_interpolation_weights.cc:32-138) --
    const int interp_method = o->interp_method;
    int interp_max = o->stencil_size;
    if (interp_max < 1) interp_max = (g.dim == 3) ? 12 : (g.dim == 2 ? 6 : 2);
    if (interp_method == FCFD_INTERP_IDW) interp_max = 99;
    {
        Wtli wtli(g.dim, interp_max, o->skewness_tol);
        g.aptr.assign(F + 1, 0); g.aidx.clear(); g.aw.clear();
        std::vector<int> st, u;
        std::vector<V3> pts;
        std::vector<double> w;
        for (int f = 0; f < F; ++f) {
            const int p = g.fparent[f], n = g.fneigh[f];
            const V3 pt = get3(g.fc, f);
            st.assign({p, n});
            if (2 < interp_max) {   // the parent's (and an interior neighbour's) neighbour cells, as a set
                u.assign(g.nb.begin() + g.nbptr[p], g.nb.begin() + g.nbptr[p + 1]);
                if (g.fbc[f] == -1) u.insert(u.end(), g.nb.begin() + g.nbptr[n], g.nb.begin() + g.nbptr[n + 1]);
                std::sort(u.begin(), u.end());
                u.erase(std::unique(u.begin(), u.end()), u.end());
                for (int x : u) if (x != p && x != n) st.push_back(x);
            }
            pts.resize(st.size());
            for (size_t q = 0; q < st.size(); ++q) pts[q] = get3(g.cc, st[q]);
            int nw;
            if (interp_method == FCFD_INTERP_WTLI) {
                nw = wtli.fit(pt, st, pts, w);
            } else if (interp_method == FCFD_INTERP_SIMPLE) {
                w.assign({0.5, 0.5});
                nw = 2;
            } else {   // "idw": distance-proportional weights (:109-118)
                w.resize(st.size());
                double sum = 0.0;
                for (size_t q = 0; q < st.size(); ++q) { w[q] = mag(pt - pts[q]); sum += w[q]; }
                for (size_t q = 0; q < st.size(); ++q) w[q] = w[q] / sum;
                nw = (int)st.size();
            }
            append_map_row(st, w, nw, g.aidx, g.aw);
            g.aptr[f + 1] = (int)g.aidx.size();
        }
    }
    // ---- node interpolation weights (node_interpolation_weights.cc:32-118) --
    // stencil: the cells around the node, boundary ghosts included, as a set
    // (ascending); the reference calls calculate_weights with is_internal =
    // true here too, so its two lowest cells always stay in a WTLI stencil
    {
        std::vector<int> ngptr(nn_local + 1, 0), nghost;
        for (int q = 0; q < B; ++q) {
            const int f = g.ghost_face[q];
            for (int k = g.fnptr[f]; k < g.fnptr[f + 1]; ++k) ngptr[g.fnodes[k] + 1]++;
        }
        for (int n = 0; n < nn_local; ++n) ngptr[n + 1] += ngptr[n];
        nghost.resize(ngptr[nn_local]);
        std::vector<int> fill(ngptr.begin(), ngptr.end() - 1);
        for (int q = 0; q < B; ++q) {
            const int f = g.ghost_face[q];
            for (int k = g.fnptr[f]; k < g.fnptr[f + 1]; ++k) nghost[fill[g.fnodes[k]]++] = N + q;
        }
        Wtli wtli(g.dim, interp_max, o->skewness_tol);
        g.naptr.assign(nn_local + 1, 0); g.naidx.clear(); g.naw.clear();
        std::vector<int> st;
        std::vector<V3> pts;
        std::vector<double> w;
        for (int n = 0; n < nn_local; ++n) {
            st.assign(ncell.begin() + ncptr[n], ncell.begin() + ncptr[n + 1]);   // ascending interior cells
            st.insert(st.end(), nghost.begin() + ngptr[n], nghost.begin() + ngptr[n + 1]);   // then ascending ghosts
            const V3 pt = get3(g.node, n);
            pts.resize(st.size());
            for (size_t q = 0; q < st.size(); ++q) pts[q] = get3(g.cc, st[q]);
            int nw;
            if (interp_method == FCFD_INTERP_WTLI) {
                nw = wtli.fit(pt, st, pts, w);
            } else if (interp_method == FCFD_INTERP_IDW) {
                w.resize(st.size());
                double sum = 0.0;
                for (size_t q = 0; q < st.size(); ++q) { w[q] = mag(pt - pts[q]); sum += w[q]; }
                for (size_t q = 0; q < st.size(); ++q) w[q] = w[q] / sum;
                nw = (int)st.size();
            } else {
                w.assign(st.size(), 1. / double(st.size()));
                nw = (int)st.size();
            }
            append_map_row(st, w, nw, g.naidx, g.naw);
            g.naptr[n + 1] = (int)g.naidx.size();
        }
    }
    // ---- gradient maps (gradient_maps.cc:39-78) ----------------------------
    g.gptr.assign(N + 1, 0); g.gidx.clear(); g.gw.clear();
    std::vector<std::pair<int, V3>> gm;
    auto gm_add = [&](int key, V3 w, int sign_if_new, int mode) {
        // mode 0: std::map::insert (no-op if present); mode 1: insert-or-accumulate with sign
        for (auto& e : gm) if (e.first == key) {
            if (mode == 1) e.second = (sign_if_new > 0) ? e.second + w : e.second - w;
            return;
        }
        gm.push_back({key, sign_if_new > 0 ? w : (-1.) * w});
    };
    for (int c = 0; c < N; ++c) {
        int nn = g.cell_nptr[c + 1] - g.cell_nptr[c];
        int meth = nn == 8 ? o->grad_hex : nn == 6 ? o->grad_prism : o->grad_other;
        gm.clear();
        V3 xc = get3(g.cc, c);
        auto other = [&](int f) { int x = g.fneigh[f]; return x == c ? g.fparent[f] : x; };
        if (meth == FCFD_GRAD_LSQR) {   // lsqr_grad_map.cc:32-99
            std::vector<int> st;
            std::vector<V3> d;
            std::vector<double> w;
            double a = 0., b = 0., cc_ = 0., dd = 0., e = 0., ff = 0.;
            auto acc = [&](int key, V3 dist) {
                double wk = dot(dist, dist);
                st.push_back(key); d.push_back(dist); w.push_back(wk);
                a += wk * dist.x * dist.x; b += wk * dist.x * dist.y; cc_ += wk * dist.x * dist.z;
                dd += wk * dist.y * dist.y; e += wk * dist.y * dist.z; ff += wk * dist.z * dist.z;
            };
            for (int k = g.nbptr[c]; k < g.nbptr[c + 1]; ++k) if (g.nb[k] != c) acc(g.nb[k], get3(g.cc, g.nb[k]) - xc);
            for (int k = g.cfptr[c]; k < g.cfptr[c + 1]; ++k) {
                int f = g.cfaces[k];
                if (g.fsym[f]) {
                    V3 nrm = get3(g.fn, f);
                    V3 point = xc + dot(get3(g.fc, f) - xc, nrm) * nrm;
                    acc(c, point - xc);
                }
            }
            double det = a * dd * ff - a * e * e - b * b * ff - cc_ * cc_ * dd + 2. * b * cc_ * e;
            for (size_t k = 0; k < st.size(); ++k) {
                V3 wt;
                wt.x = w[k] * (d[k].x * (dd * ff - e * e) + d[k].y * (cc_ * e - b * ff) + d[k].z * (b * e - cc_ * dd));
                wt.y = w[k] * (d[k].x * (cc_ * e - b * ff) + d[k].y * (a * ff - cc_ * cc_) + d[k].z * (b * cc_ - a * e));
                wt.z = w[k] * (d[k].x * (b * e - cc_ * dd) + d[k].y * (b * cc_ - a * e) + d[k].z * (a * dd - b * b));
                wt = wt / det;
                gm_add(st[k], wt, +1, 0);   // map insert: no-op when the key exists
                gm_add(c, wt, -1, 1);       // self weight -= w
            }
        } else if (meth == FCFD_GRAD_CURVILINEAR) {   // curvilinear_grad_map.cc:28-228
            int cp[6] = {-1, -1, -1, -1, -1, -1};
            V3 J[3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            int nf = g.cfptr[c + 1] - g.cfptr[c];
            const int* cf = &g.cfaces[g.cfptr[c]];
            if (nn == 8) {
                // greedy pairing of the remaining slots (intended semantics of :57-78)
                int rem[8], nrem = nf, pairs[6] = {-1, -1, -1, -1, -1, -1};
                for (int q = 0; q < nf; ++q) rem[q] = q;
                for (int cnt = 0; cnt < 3 && nrem > 0; ++cnt) {
                    int s1 = rem[0];
                    int f1 = cf[s1];
                    V3 my = get3(g.fn, f1);
                    if (g.fparent[f1] != c) my = {my.x * -1., my.y * -1., my.z * -1.};
                    pairs[2 * cnt + 1] = s1;
                    double mn = 1.e20;
                    for (int q = 1; q < nrem; ++q) {
                        int f2 = cf[rem[q]];
                        V3 ot = get3(g.fn, f2);
                        if (g.fparent[f2] != c) ot = {ot.x * -1., ot.y * -1., ot.z * -1.};
                        double pr = dot(my, ot);
                        if (pr < mn) { mn = pr; pairs[2 * cnt] = rem[q]; }
                    }
                    int k2 = 0;
                    for (int q = 0; q < nrem; ++q) if (rem[q] != pairs[2 * cnt] && rem[q] != pairs[2 * cnt + 1]) rem[k2++] = rem[q];
                    nrem = k2;
                }
                for (int i = 0; i < 3; ++i) {
                    cp[2 * i + 1] = other(cf[pairs[2 * i + 1]]);
                    cp[2 * i] = other(cf[pairs[2 * i]]);
                    J[i] = get3(g.cc, cp[2 * i + 1]) - get3(g.cc, cp[2 * i]);
                }
            } else {
                int st[8];
                for (int q = 0; q < nf; ++q) st[q] = cf[q];   // already ascending
                double maxdet = 0.0;
                for (int i1 = 0; i1 < nf; ++i1)
                    for (int i2 = i1; i2 < nf; ++i2) {
                        int p1 = other(st[i2]), m1 = other(st[i1]);
                        V3 j0 = get3(g.cc, p1) - get3(g.cc, m1);
                        for (int i3 = i1; i3 < nf; ++i3)
                            for (int i4 = i3; i4 < nf; ++i4) {
                                int p2 = other(st[i4]), m2 = other(st[i3]);
                                V3 j1 = get3(g.cc, p2) - get3(g.cc, m2);
                                for (int i5 = i2; i5 < nf; ++i5)
                                    for (int i6 = i5; i6 < nf; ++i6) {
                                        int p3 = other(st[i6]), m3 = other(st[i5]);
                                        V3 j2 = get3(g.cc, p3) - get3(g.cc, m3);
                                        double det = dot(j0, cross(j1, j2));
                                        if (det > maxdet) {
                                            maxdet = det;
                                            cp[0] = m1; cp[1] = p1; cp[2] = m2; cp[3] = p2; cp[4] = m3; cp[5] = p3;
                                            J[0] = j0; J[1] = j1; J[2] = j2;
                                        }
                                    }
                            }
                    }
            }
            double det = dot(J[0], cross(J[1], J[2]));
            V3 Ji[3];
            Ji[0].x = (J[1].y * J[2].z - J[1].z * J[2].y) / det;
            Ji[1].x = (J[0].z * J[2].y - J[0].y * J[2].z) / det;
            Ji[2].x = (J[0].y * J[1].z - J[0].z * J[1].y) / det;
            Ji[0].y = (J[1].z * J[2].x - J[1].x * J[2].z) / det;
            Ji[1].y = (J[0].x * J[2].z - J[0].z * J[2].x) / det;
            Ji[2].y = (J[0].z * J[1].x - J[0].x * J[1].z) / det;
            Ji[0].z = (J[1].x * J[2].y - J[1].y * J[2].x) / det;
            Ji[1].z = (J[0].y * J[2].x - J[0].x * J[2].y) / det;
            Ji[2].z = (J[0].x * J[1].y - J[0].y * J[1].x) / det;
            gm.push_back({cp[0], (-1.) * Ji[0]});
            for (int q = 1; q < 6; ++q) gm_add(cp[q], Ji[q / 2], (q & 1) ? +1 : -1, 1);
        }
        std::sort(gm.begin(), gm.end(), [](const std::pair<int, V3>& x, const std::pair<int, V3>& y) { return x.first < y.first; });
        for (auto& e : gm) { g.gidx.push_back(e.first); g.gw.push_back(e.second.x); g.gw.push_back(e.second.y); g.gw.push_back(e.second.z); }
        g.gptr[c + 1] = (int)g.gidx.size();
    }
    g.perm.resize(T); g.iperm.resize(T); g.fperm.resize(F); g.ifperm.resize(F);
    std::iota(g.perm.begin(), g.perm.end(), 0);
    std::iota(g.iperm.begin(), g.iperm.end(), 0);
    std::iota(g.fperm.begin(), g.fperm.end(), 0);
    std::iota(g.ifperm.begin(), g.ifperm.end(), 0);
    return FCFD_OK;
}

// Reverse Cuthill-McKee over the interior face graph: small bandwidth, but the
// neighbours of a cell sit about one BFS level (a whole cross-section) ahead.
static std::vector<int> rcm_order(const HostGrid& g) {
    const int N = g.N, F = g.F;
    std::vector<int> deg(N, 0);
    for (int f = 0; f < F; ++f) if (g.fbc[f] < 0) { deg[g.fparent[f]]++; deg[g.fneigh[f]]++; }
    std::vector<int> ptr(N + 1, 0), adj;
    for (int c = 0; c < N; ++c) ptr[c + 1] = ptr[c] + deg[c];
    adj.resize(ptr[N]);
    {
        std::vector<int> fill(ptr.begin(), ptr.end() - 1);
        for (int f = 0; f < F; ++f) if (g.fbc[f] < 0) { adj[fill[g.fparent[f]]++] = g.fneigh[f]; adj[fill[g.fneigh[f]]++] = g.fparent[f]; }
    }
    std::vector<int> order;
    order.reserve(N);
    std::vector<char> seen(N, 0);
    std::vector<int> nbrs, q, mark(N, -1);
    int sweep = 0;
    for (int start = 0; start < N; ++start) {
        if (seen[start]) continue;
        // pseudo-peripheral seed: two breadth-first sweeps, each restarting from
        // the last vertex reached by the previous one
        int s = start;
        for (int it = 0; it < 2; ++it, ++sweep) {
            q.assign(1, s);
            mark[s] = sweep;
            size_t h = 0;
            while (h < q.size()) {
                int x = q[h++];
                for (int k = ptr[x]; k < ptr[x + 1]; ++k) if (mark[adj[k]] != sweep) { mark[adj[k]] = sweep; q.push_back(adj[k]); }
            }
            s = q.back();
        }
        size_t head = order.size();
        order.push_back(s);
        seen[s] = 1;
        while (head < order.size()) {
            int x = order[head++];
            nbrs.clear();
            for (int k = ptr[x]; k < ptr[x + 1]; ++k) if (!seen[adj[k]]) { seen[adj[k]] = 1; nbrs.push_back(adj[k]); }
            std::sort(nbrs.begin(), nbrs.end(), [&](int a, int b) { return deg[a] != deg[b] ? deg[a] < deg[b] : a < b; });
            order.insert(order.end(), nbrs.begin(), nbrs.end());
        }
    }
    std::reverse(order.begin(), order.end());
    return order;
}

// Space-filling-curve order: cells sorted by the Morton code of their centroid
// in rank space.  Along each axis a cell's coordinate is replaced by the number
// of cells with a strictly smaller centroid coordinate, scaled to 21 bits, so
// that stretched and clustered meshes are spread evenly (on a structured block
// the ranks are the block indices times a constant); the codes interleave the
// three ranks, ties broken by reference index.  Runs of consecutive cells are
// then compact blobs, so the cells a face or cell gathers mostly lie in the
// cache lines the same or the previous few wavefronts already fetched.
static std::vector<int> sfc_order(const HostGrid& g) {
    const int N = g.N;
    std::vector<uint64_t> code(N, 0);
    std::vector<int> idx(N);
    for (int d = 0; d < 3; ++d) {
        std::iota(idx.begin(), idx.end(), 0);
        std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return g.cc[3 * a + d] < g.cc[3 * b + d]; });
        int rank = 0;
        for (int k = 0; k < N; ++k) {
            if (k > 0 && g.cc[3 * idx[k] + d] != g.cc[3 * idx[k - 1] + d]) rank = k;
            uint64_t q = (uint64_t)(((unsigned __int128)rank << 21) / (unsigned)N);   // < 2^21
            uint64_t& c = code[idx[k]];
            for (int b = 0; b < 21; ++b) c |= ((q >> b) & 1ull) << (3 * b + d);
        }
    }
    std::vector<int> order(N);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return code[a] < code[b]; });
    return order;
}

// Internal numbering of cells (RCM or space-filling curve); ghosts follow in
// the order of their (renumbered) parents; faces sorted by renumbered parent.
void renumber(HostGrid& g, int cell_order) {
    const int N = g.N, B = g.B, F = g.F;
    std::vector<int> order = cell_order == kOrderRcm ? rcm_order(g) : sfc_order(g);
    for (int i = 0; i < N; ++i) { g.iperm[i] = order[i]; g.perm[order[i]] = i; }
    // ghosts: ordered by renumbered parent, then reference face order
    std::vector<int> gh(B);
    std::iota(gh.begin(), gh.end(), 0);
    std::stable_sort(gh.begin(), gh.end(), [&](int a, int b) { return g.perm[g.fparent[g.ghost_face[a]]] < g.perm[g.fparent[g.ghost_face[b]]]; });
    for (int i = 0; i < B; ++i) { g.iperm[N + i] = N + gh[i]; g.perm[N + gh[i]] = N + i; }
    // faces: by renumbered parent, then reference order
    std::vector<int> fo(F);
    std::iota(fo.begin(), fo.end(), 0);
    std::stable_sort(fo.begin(), fo.end(), [&](int a, int b) { return g.perm[g.fparent[a]] < g.perm[g.fparent[b]]; });
    for (int i = 0; i < F; ++i) { g.ifperm[i] = fo[i]; g.fperm[fo[i]] = i; }
}

}  // namespace fcfd
