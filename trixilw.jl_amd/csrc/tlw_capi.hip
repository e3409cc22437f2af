// tlw_capi.hip -- C ABI of libtrixilw_hip.so (declared in include/trixilw_hip.h).
//
// Host side of the drop-in boundary: validates and uploads the mesh/basis
// description (what SemidiscretizationHyperbolic + create_cache_serial build,
// src/semidiscretization/semidiscretization_hyperbolic.jl:58-84,
// src/solvers/dgsem_tree/containers.jl:7-64), owns the device buffers, and
// sequences the kernels of tlw_device.hpp for rhs!, rhs_mdrk1!/2!, max_dt and
// the device-resident solve_lwfr loop.  A context holds one mesh or a batch of
// independent meshes of the same polynomial degree (instances), each with its
// own time, dt and CFL reduction, advanced by the same launches.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/trixilw_hip.h"
#include "tlw_adjoint.hpp"
#include "tlw_device.hpp"
#include "tlw_noalloc.hpp"

using namespace tlw;

struct tlw_ctx;
// Two batch contexts, one of polydeg 3 and one of polydeg 4, advanced by the
// same launches (tlw_create_ensemble).
struct tlw_ensemble {
    tlw_ctx* c3 = nullptr;
    tlw_ctx* c4 = nullptr;
    hipStream_t stream = nullptr;
    bool graphs = false;
    std::map<int, hipGraphExec_t> graph_cache;
    hipStream_t graph_stream = nullptr;
    std::string err;
};

struct tlw_ctx {
    int N = 0, n = 0, E = 0;
    int ninst = 1;
    std::vector<int> inst_first;  // [ninst + 1] first element of each instance
    int time_disc = TLW_LW;
    int flags = 0;
    hipStream_t stream = nullptr;
    DevConst hconst{};
    DevConst* d_const = nullptr;
    DevState* d_state = nullptr;  // [ninst]
    DevState* h_state = nullptr;  // pinned mirror [ninst]
    int* d_cur = nullptr;         // [ninst] staging for k_init_state
    int* d_nbr = nullptr;
    double* d_invj = nullptr;
    double* d_center = nullptr;
    double* d_length = nullptr;
    double* d_trace = nullptr;
    double* d_buf[2] = {nullptr, nullptr};
    double* d_us = nullptr;
    double* d_auxU = nullptr;
    double* d_auxF = nullptr;
    double* d_partial = nullptr;
    double* d_alpha_tmp = nullptr;
    double* d_alpha = nullptr;
    // workgroup tables (KArgs::vblk, sblk, etile) and per-instance partial ranges
    int4* d_vblk = nullptr;
    int4* d_sblk = nullptr;
    int* d_etile = nullptr;
    int4* d_rng = nullptr;
    int vgrid = 0, sgrid = 0, vcap = 0;
    bool sc = false;
    bool parabolic = false;   // compressible Navier-Stokes (tlw_create_parabolic)
    // curvilinear StructuredMesh (tlw_create_curved): per node Ja [E][n²][4] and
    // J⁻¹ [E][n²], per face node the interface normal (n1, n2, |n|, 0) [E][4][n][4]
    bool curved = false;
    double* d_Ja = nullptr;
    double* d_ijn = nullptr;
    double* d_fnorm = nullptr;
    double* d_fxy = nullptr;  // boundary face-node coordinates (curvilinear mesh with boundaries)
    double* d_nxy = nullptr;  // node coordinates [E][n*n][2] (curvilinear mesh with source terms)
    double* d_fvb = nullptr;   // Navier-Stokes walls: viscous time average of boundary face nodes
    double* d_snrm = nullptr;  // subcell normals (curvilinear shock capturing) [E][n²][4]
    double* d_sqrk = nullptr;  // calc_fn_low! normals of subcell face n-1 [E][n][4]
    bool bad_launch = false;  // a launch was refused for a missing input (reported by check_launch)
    size_t slot_off = 0;      // offset of the CFL maxima slots in d_partial
    int pv_off = 0;           // offset of the viscous CFL partials (0: none)
    int ebv = 0;              // elements per volume workgroup
    bool positivity = false;  // Zhang-Shu limiter in the device loop
    bool adaptive = false;    // error-controlled steps (tlw_solve_init_adaptive)
    bool has_boundary = false;  // a face with neighbour -1, or a mortar (paths that refuse both test this)
    bool bc_faces = false;      // a face with neighbour -1
    // L2 mortars of a TreeMesh with refined cells: per mortar (orientation,
    // large_side), the volume kernel's face-node values and the mortar fluxes
    int n_mortars = 0;
    // P4estMesh: interface partner per element face (KArgs::plink)
    bool p4est = false;
    int4* d_plink = nullptr;
    int2* d_mort = nullptr;
    double* d_mrec = nullptr;
    double* d_mflux = nullptr;
    std::vector<int> host_cur;  // [ninst] state buffer holding each instance's uⁿ
    bool aux_valid = false;
    bool alpha_valid = false;  // blending factors of the last rhs or step (shock capturing)
    // kernel timing (tlw_enable_kernel_timing)
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, int>> ev_pairs;  // (start index, kind 0 volume / 1 surface / 2 limiter)
    size_t ev_used = 0;
    double t_ms[3] = {0.0, 0.0, 0.0};
    int64_t t_n[3] = {0, 0, 0};
    // captured step sequences of tlw_advance, by step count (tlw_enable_graphs)
    // caller-defined Dirichlet data (TLW_BV_TABLE): boundary face index per element face,
    // face-node coordinates in table order, the uploaded table and its time levels (0: none yet)
    int* d_bidx = nullptr;
    long long n_bnodes = 0;
    std::vector<double> bxy;
    double* d_btab = nullptr;
    size_t btab_cap = 0;
    int btab_levels = 0;
    double* d_adj_tmp = nullptr;    // the MDRK adjoint's λ_us (allocated on first use)
    double* d_adj_trace = nullptr;  // face records of the blended adjoint (allocated on first use)
    // per-step dt record and prescribed dts of the device loop (tlw_record_dts, tlw_prescribe_dts)
    double* d_dtlog = nullptr;
    long long dtlog_cap = 0;
    double* d_dtpresc = nullptr;
    long long dtpresc_n = 0;
    bool graphs = false;
    std::map<int, hipGraphExec_t> graph_cache;
    hipStream_t graph_stream = nullptr;
    std::string err;
};

namespace {

int fail(tlw_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(ctx, call)                                                                         \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail((ctx), TLW_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));      \
    } while (0)

size_t state_len(const tlw_ctx* c) { return (size_t)4 * c->n * c->n * c->E; }
size_t inst_offset(const tlw_ctx* c, int k) { return (size_t)4 * c->n * c->n * c->inst_first[k]; }
size_t inst_len(const tlw_ctx* c, int k) {
    return (size_t)4 * c->n * c->n * (c->inst_first[k + 1] - c->inst_first[k]);
}

// Elements per volume workgroup: the compiled tile sizes, selectable at run
// time (tlw_set_elements_per_workgroup, tlw_autotune).  polydeg 4 (25 nodes):
// 5, 10, 15 elements = 125, 250, 375 lanes of 128, 256, 384; polydeg 3
// (16 nodes): 4, 8, 16 elements = 64, 128, 256 lanes, no idle lanes.
constexpr int kEbChoices4[3] = {5, 10, 15};
constexpr int kEbChoices3[3] = {4, 8, 16};
const int* eb_choices(int N) { return N == 4 ? kEbChoices4 : kEbChoices3; }
int default_eb(int N) { return N == 4 ? 15 : 16; }
int min_eb(int N) { return N == 4 ? 5 : 4; }
bool eb_supported(int N, int eb) {
    const int* ch = eb_choices(N);
    return eb == ch[0] || eb == ch[1] || eb == ch[2];
}
int surf_eb(int N) { return N == 4 ? SGeo<4>::EB : SGeo<3>::EB; }

void drop_graphs(tlw_ctx* c) {
    for (auto& kv : c->graph_cache) (void)hipGraphExecDestroy(kv.second);
    c->graph_cache.clear();
}

void free_mesh(tlw_ctx* c) {
    drop_graphs(c);
    void* ptrs[] = {c->d_nbr,       c->d_invj,  c->d_center, c->d_length, c->d_trace,   c->d_buf[0],    c->d_buf[1],
                    c->d_us,        c->d_auxU,  c->d_auxF,   c->d_partial, c->d_alpha_tmp, c->d_alpha,
                    c->d_vblk,      c->d_sblk,  c->d_etile,  c->d_rng,    c->d_state,   c->d_cur,
                    c->d_Ja,        c->d_ijn,   c->d_fnorm,  c->d_fxy,     c->d_snrm,      c->d_sqrk,
                    c->d_fvb,       c->d_bidx,  c->d_btab,   c->d_mort,   c->d_mrec,    c->d_mflux,
                    c->d_plink,     c->d_nxy};
    for (void* p : ptrs) (void)hipFree(p);
    if (c->h_state) (void)hipHostFree(c->h_state);
    c->d_nbr = nullptr;
    c->d_invj = c->d_center = c->d_length = c->d_trace = c->d_us = c->d_auxU = c->d_auxF = nullptr;
    c->d_buf[0] = c->d_buf[1] = nullptr;
    c->d_partial = c->d_alpha_tmp = c->d_alpha = nullptr;
    c->d_vblk = c->d_sblk = c->d_rng = nullptr;
    c->d_etile = nullptr;
    c->d_state = nullptr;
    c->h_state = nullptr;
    c->d_cur = nullptr;
    c->d_Ja = c->d_ijn = c->d_fnorm = c->d_fxy = c->d_snrm = c->d_sqrk = c->d_fvb = nullptr;
    c->d_bidx = nullptr;
    c->d_btab = nullptr;
    c->d_mort = nullptr;
    c->d_mrec = c->d_mflux = nullptr;
    c->d_plink = nullptr;
    c->d_nxy = nullptr;
    c->n_mortars = 0;
    c->btab_cap = 0;
    c->btab_levels = 0;
    c->n_bnodes = 0;
    c->bxy.clear();
}

// Workgroup tables for the current tile: volume workgroups of ebv elements and
// node-per-lane workgroups of the surface size, both restarted at every
// instance boundary, the element -> volume workgroup map and each instance's
// range of partials (volume workgroups, then the node-per-lane grid after them).
int build_tables(tlw_ctx* c) {
    std::vector<int4> vb, sb, rng(c->ninst);
    std::vector<int> et(c->E);
    const int seb = surf_eb(c->N);
    for (int k = 0; k < c->ninst; ++k) {
        const int f0 = c->inst_first[k], f1 = c->inst_first[k + 1];
        rng[k].x = (int)vb.size();
        for (int e = f0; e < f1; e += c->ebv) {
            const int nel = std::min(c->ebv, f1 - e);
            for (int q = 0; q < nel; ++q) et[e + q] = (int)vb.size();
            vb.push_back(make_int4(e, nel, k, 0));
        }
        rng[k].y = (int)vb.size();
        rng[k].z = (int)sb.size();
        for (int e = f0; e < f1; e += seb) sb.push_back(make_int4(e, std::min(seb, f1 - e), k, 0));
        rng[k].w = (int)sb.size();
    }
    if ((int)vb.size() > c->vcap) return fail(c, TLW_EINVAL, "volume workgroup table exceeds its allocation");
    c->vgrid = (int)vb.size();
    c->sgrid = (int)sb.size();
    for (auto& r : rng) {  // surface partials follow the volume ones
        r.z += c->vgrid;
        r.w += c->vgrid;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(c->d_vblk, vb.data(), sizeof(int4) * vb.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_sblk, sb.data(), sizeof(int4) * sb.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_etile, et.data(), sizeof(int) * et.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_rng, rng.data(), sizeof(int4) * rng.size(), hipMemcpyHostToDevice));
    drop_graphs(c);
    return TLW_OK;
}

int validate_mesh(tlw_ctx* c, const tlw_mesh_desc* m, bool mortars = false) {
    if (!m || m->n_elements <= 0 || !m->neighbors_host)
        return fail(c, TLW_EINVAL, "mesh description: missing neighbour table or n_elements <= 0");
    if (!c->curved && (!m->inverse_jacobian_host || !m->element_center_host || !m->element_length_host))
        return fail(c, TLW_EINVAL, "mesh description: missing arrays");
    const int E = m->n_elements;
    const int lowest = mortars ? TLW_NEIGHBOR_MORTAR : -1;
    for (size_t k = 0; k < (size_t)4 * E; ++k) {
        const int nb = m->neighbors_host[k];
        if (nb < lowest || nb >= E) return fail(c, TLW_EINVAL, "mesh description: neighbour id out of range");
    }
    // conformity: the neighbour across a face must point back through the opposite face
    // (a P4estMesh's faces meet in any orientation: its interface list is checked instead)
    for (int e = 0; e < E && !c->p4est; ++e)
        for (int f = 0; f < 4; ++f) {
            const int nb = m->neighbors_host[4 * e + f];
            if (nb >= 0 && m->neighbors_host[4 * nb + (f ^ 1)] != e)
                return fail(c, TLW_EINVAL, "mesh description: neighbour table is not symmetric (non-conforming mesh)");
        }
    if (c->curved) return TLW_OK;  // metrics are checked with the curved description
    for (int e = 0; e < E; ++e)
        if (!(m->inverse_jacobian_host[e] > 0.0) || !(m->element_length_host[e] > 0.0))
            return fail(c, TLW_EINVAL, "mesh description: non-positive Jacobian or element length");
    return TLW_OK;
}

// Curvilinear metrics: finite contravariant vectors, positive inverse Jacobians
// (the normal's sign factor of calc_interface_flux! is then 1), and the
// interface normals of every face node, taken from the right element's face
// node so that both elements of an interface use the same bits.
// Subcell normals of Trixi's calcflux_fv! for curvilinear meshes: direction d
// (1: ξ, 2: η), row r of element e, subcell face k: Ja^d of the row's first node
// plus, for each subcell q < k, Σ_m w_q D[q][m] Ja^d(m), summed in that order;
// with quirk, the q = 0 and q = n-2 terms only (the reference's calc_fn_low!,
// structured/dg_2d.jl:111-173, loops over the tuple (2, nnodes)).
void subcell_normal(const DevConst& h, int n, const double* Ja, int d, size_t e, int r, int k, bool quirk,
                    double* nrm) {
    const size_t NN = (size_t)n * n;
    auto ja = [&](int m) {
        const size_t node = e * NN + (d == 1 ? (size_t)n * r + m : (size_t)n * m + r);
        return Ja + 4 * node + (d == 1 ? 0 : 2);
    };
    nrm[0] = ja(0)[0];
    nrm[1] = ja(0)[1];
    for (int q = 0; q < k; ++q) {
        if (quirk && q != 0 && q != n - 2) continue;
        for (int m = 0; m < n; ++m) {
            const double cq = h.weights[q] * h.D[q][m];
            nrm[0] = nrm[0] + cq * ja(m)[0];
            nrm[1] = nrm[1] + cq * ja(m)[1];
        }
    }
}

int prepare_curved(tlw_ctx* c, const tlw_mesh_desc* m, const tlw_curved_desc* cv, std::vector<double>& fnorm,
                   std::vector<double>& invj, std::vector<double>& fxy) {
    if (!cv || !cv->contravariant_vectors_host || !cv->inverse_jacobian_host)
        return fail(c, TLW_EINVAL, "curved description: missing arrays");
    const int E = m->n_elements, n = c->n, NN = n * n;
    const double* Ja = cv->contravariant_vectors_host;
    const double* ijn = cv->inverse_jacobian_host;
    for (size_t k = 0; k < (size_t)4 * NN * E; ++k)
        if (!std::isfinite(Ja[k])) return fail(c, TLW_EINVAL, "curved description: non-finite contravariant vector");
    for (size_t k = 0; k < (size_t)NN * E; ++k)
        if (!(ijn[k] > 0.0) || !std::isfinite(ijn[k]))
            return fail(c, TLW_EINVAL, "curved description: inverse Jacobian must be finite and > 0");
    bool boundary = false;
    for (size_t k = 0; k < (size_t)4 * E; ++k) boundary = boundary || m->neighbors_host[k] < 0;
    const double* xy = cv->node_coordinates_host;
    if (boundary && !xy) return fail(c, TLW_EINVAL, "curved description: node coordinates needed on a mesh with boundaries");
    if (c->hconst.src_kind != TLW_SRC_NONE && !xy)
        return fail(c, TLW_EINVAL, "curved description: node coordinates needed with source terms");
    fnorm.assign((size_t)E * 4 * n * 4, 0.0);
    fxy.assign(boundary ? (size_t)E * 4 * n * 2 : 0, 0.0);
    invj.resize(E);
    for (int e = 0; e < E; ++e) {
        invj[e] = ijn[(size_t)e * NN];
        for (int f = 0; f < 4; ++f) {
            const int nb = m->neighbors_host[4 * e + f];
            // an interface takes the right element's face node; a domain boundary the element's own;
            // on a P4estMesh every face its own (each side of an interface uses its own normal)
            const int right = c->p4est ? e : (f & 1) ? (nb >= 0 ? nb : e) : e;
            const bool own_plus = (f & 1) && (nb < 0 || c->p4est);
            for (int l = 0; l < n; ++l) {
                // x faces: Ja¹ at node (1, l) of the right element (own: (n, l)); y faces: Ja² at (l, 1) ((l, n))
                const size_t loc = f < 2 ? (size_t)n * l + (own_plus ? n - 1 : 0)
                                         : (size_t)l + (own_plus ? (size_t)n * (n - 1) : 0);
                const size_t node = (size_t)right * NN + loc;
                if (nb < 0) {
                    fxy[((((size_t)e * 4 + f) * n + l) * 2)] = xy[2 * node];
                    fxy[((((size_t)e * 4 + f) * n + l) * 2) + 1] = xy[2 * node + 1];
                }
                const double n1 = Ja[4 * node + (f < 2 ? 0 : 2)], n2 = Ja[4 * node + (f < 2 ? 1 : 3)];
                double* o = &fnorm[(((size_t)e * 4 + f) * n + l) * 4];
                o[0] = n1;
                o[1] = n2;
                o[2] = std::sqrt(n1 * n1 + n2 * n2);
            }
        }
    }
    return TLW_OK;
}

// Uploads ninst meshes as one element range each (element ids offset by the
// instance's first element) and allocates everything that depends on E.
// Mortar list of a TreeMesh with refined cells: every face marked
// TLW_NEIGHBOR_MORTAR lies on exactly one mortar, in the role the record gives
// it, and the geometry is a 2:1 face (small elements of half the large one's
// length, side by side along the large face, lower before upper).  Returns the
// device neighbour code of each mortar face in code[4*E] (-2 - (3 m + role)).
int validate_mortars(tlw_ctx* c, const tlw_mesh_desc* m, const tlw_mortar_desc* md, std::vector<int>& code) {
    const int E = m->n_elements, M = md->n_mortars;
    if (M < 0 || (M > 0 && !md->mortars_host)) return fail(c, TLW_EINVAL, "mortar description: missing records");
    code.assign(4 * (size_t)E, 0);
    const double* xc = m->element_center_host;
    const double* len = m->element_length_host;
    for (int q = 0; q < M; ++q) {
        const int* r = md->mortars_host + 5 * (size_t)q;
        const int lower = r[0], upper = r[1], large = r[2], o = r[3], side = r[4];
        if (lower < 0 || lower >= E || upper < 0 || upper >= E || large < 0 || large >= E || lower == upper ||
            (o != 1 && o != 2) || (side != 1 && side != 2))
            return fail(c, TLW_EINVAL, "mortar description: element id, orientation or large side out of range");
        const int lface = (side == 1) ? 2 * o - 1 : 2 * o - 2, sface = lface ^ 1;
        const int roles[3][2] = {{lower, sface}, {upper, sface}, {large, lface}};
        for (int role = 0; role < 3; ++role) {
            const size_t k = 4 * (size_t)roles[role][0] + roles[role][1];
            if (m->neighbors_host[k] != TLW_NEIGHBOR_MORTAR || code[k] != 0)
                return fail(c, TLW_EINVAL, "mortar description: a mortar face is not marked or lies on two mortars");
            code[k] = -2 - (3 * q + role);
        }
        // geometry: half-size small elements next to each other, lower first along the face
        const int ax = o - 1, tx = 1 - ax;
        const double h = len[large], tol = 1e-9 * h;
        if (std::fabs(len[lower] - 0.5 * h) > tol || std::fabs(len[upper] - 0.5 * h) > tol ||
            std::fabs((xc[2 * upper + tx] - xc[2 * lower + tx]) - 0.5 * h) > tol ||
            std::fabs(0.5 * (xc[2 * upper + tx] + xc[2 * lower + tx]) - xc[2 * large + tx]) > tol)
            return fail(c, TLW_EINVAL, "mortar description: the small elements do not tile the large face");
        (void)ax;
    }
    for (size_t k = 0; k < 4 * (size_t)E; ++k)
        if (m->neighbors_host[k] == TLW_NEIGHBOR_MORTAR && code[k] == 0)
            return fail(c, TLW_EINVAL, "mortar description: a face marked TLW_NEIGHBOR_MORTAR has no mortar");
    return TLW_OK;
}

int upload_meshes(tlw_ctx* c, const tlw_mesh_desc* ms, int ninst, const tlw_curved_desc* cv = nullptr,
                  const tlw_mortar_desc* md = nullptr, const tlw_p4est_desc* pd = nullptr) {
    if (ninst < 1) return fail(c, TLW_EINVAL, "at least one instance");
    if (c->curved && ninst != 1) return fail(c, TLW_ENOTSUP, "a curvilinear mesh is a single instance");
    long long total = 0;
    for (int k = 0; k < ninst; ++k) {
        int rc = validate_mesh(c, &ms[k], md != nullptr);
        if (rc) return rc;
        total += ms[k].n_elements;
    }
    if (total > (1LL << 30)) return fail(c, TLW_EINVAL, "too many elements");
    std::vector<int> mcode;
    const int M = md ? md->n_mortars : 0;
    if (md) {
        if (ninst != 1 || c->curved || c->sc || c->parabolic)
            return fail(c, TLW_ENOTSUP, "mortars: a single-instance hyperbolic TreeMesh context with VolumeIntegralFR");
        int rc = validate_mortars(c, &ms[0], md, mcode);
        if (rc) return rc;
    }
    std::vector<int4> plink;
    if (pd) {
        const int E0 = ms[0].n_elements, I = pd->n_interfaces;
        if (ninst != 1 || I < 0 || (I > 0 && !pd->interfaces_host))
            return fail(c, TLW_EINVAL, "P4est description: one instance and an interface list");
        plink.assign(4 * (size_t)E0, make_int4(-1, -1, 0, 0));
        for (int k = 0; k < I; ++k) {
            const int* r = pd->interfaces_host + 5 * (size_t)k;
            const int pe = r[0], pf = r[1], se = r[2], sf = r[3], fl = r[4];
            if (pe < 0 || pe >= E0 || se < 0 || se >= E0 || pf < 0 || pf > 3 || sf < 0 || sf > 3 || (fl != 0 && fl != 1) ||
                (pe == se && pf == sf))
                return fail(c, TLW_EINVAL, "P4est description: interface entry out of range");
            int4& a = plink[4 * (size_t)pe + pf];
            int4& b = plink[4 * (size_t)se + sf];
            if (a.x >= 0 || b.x >= 0) return fail(c, TLW_EINVAL, "P4est description: a face lies on two interfaces");
            if (ms[0].neighbors_host[4 * (size_t)pe + pf] != se || ms[0].neighbors_host[4 * (size_t)se + sf] != pe)
                return fail(c, TLW_EINVAL, "P4est description: neighbours_host disagrees with the interface list");
            a = make_int4(se, sf, fl, 1);
            b = make_int4(pe, pf, fl, 0);
        }
        for (size_t q = 0; q < 4 * (size_t)E0; ++q)
            if (plink[q].x < 0 && ms[0].neighbors_host[q] != -1)
                return fail(c, TLW_EINVAL, "P4est description: a face with a neighbour is on no interface");
    }
    bool boundary = false;
    for (int k = 0; k < ninst; ++k)
        for (size_t q = 0; q < (size_t)4 * ms[k].n_elements; ++q)
            if (ms[k].neighbors_host[q] == -1) boundary = true;
    if (boundary && c->hconst.bc_kind == TLW_BV_NONE)
        return fail(c, TLW_EINVAL, "mesh has boundary faces but no boundary value function was given");
    if (boundary && c->parabolic && !c->hconst.ns_walls)
        return fail(c, TLW_ENOTSUP, "the hyperbolic-parabolic path needs a periodic mesh or walls");
    std::vector<double> fnorm, cinvj, fxy;
    std::vector<double> snrm, sqrk;
    if (c->curved) {
        int rc = prepare_curved(c, &ms[0], cv, fnorm, cinvj, fxy);
        if (rc) return rc;
        if (c->sc) {
            const int n = c->n, NN = n * n, Ec = ms[0].n_elements;
            snrm.assign((size_t)Ec * NN * 4, 0.0);
            sqrk.assign((size_t)Ec * n * 4, 0.0);
            const double* Ja = cv->contravariant_vectors_host;
            for (size_t e = 0; e < (size_t)Ec; ++e) {
                for (int jj = 0; jj < n; ++jj)
                    for (int ii = 0; ii < n; ++ii) {
                        double* o = &snrm[(e * NN + (size_t)ii + (size_t)n * jj) * 4];
                        if (ii > 0) subcell_normal(c->hconst, n, Ja, 1, e, jj, ii, false, o);
                        if (jj > 0) subcell_normal(c->hconst, n, Ja, 2, e, ii, jj, false, o + 2);
                    }
                for (int l = 0; l < n; ++l) {
                    subcell_normal(c->hconst, n, Ja, 1, e, l, n - 1, true, &sqrk[(e * n + l) * 4]);
                    subcell_normal(c->hconst, n, Ja, 2, e, l, n - 1, true, &sqrk[(e * n + l) * 4 + 2]);
                }
            }
        }
    }
    free_mesh(c);
    c->aux_valid = c->alpha_valid = false;
    c->ninst = ninst;
    c->E = (int)total;
    // the P4est interfaces are paired in the surface kernel's boundary-capable path
    c->has_boundary = boundary || M > 0 || pd != nullptr;
    c->bc_faces = boundary;
    c->n_mortars = M;
    c->inst_first.assign(ninst + 1, 0);
    for (int k = 0; k < ninst; ++k) c->inst_first[k + 1] = c->inst_first[k] + ms[k].n_elements;
    c->host_cur.assign(ninst, 0);
    const int E = c->E;
    std::vector<int> nbr(4 * (size_t)E);
    std::vector<double> invj(E), center(2 * (size_t)E), length(E);
    for (int k = 0; k < ninst; ++k) {
        const int f0 = c->inst_first[k], Ek = ms[k].n_elements;
        for (size_t q = 0; q < (size_t)4 * Ek; ++q) {
            const int nb = ms[k].neighbors_host[q];
            nbr[4 * (size_t)f0 + q] = pd ? (nb < 0 ? -1 : -3) : nb == TLW_NEIGHBOR_MORTAR ? mcode[q] : (nb < 0 ? -1 : nb + f0);
        }
        if (c->curved) {  // element data of a TreeMesh: unused, J⁻¹ of each element's first node
            std::memcpy(&invj[f0], cinvj.data(), sizeof(double) * Ek);
            continue;
        }
        std::memcpy(&invj[f0], ms[k].inverse_jacobian_host, sizeof(double) * Ek);
        std::memcpy(&center[2 * (size_t)f0], ms[k].element_center_host, sizeof(double) * 2 * Ek);
        std::memcpy(&length[f0], ms[k].element_length_host, sizeof(double) * Ek);
    }
    const size_t nd = state_len(c);
    HIPCHK(c, hipMalloc(&c->d_nbr, sizeof(int) * 4 * E));
    HIPCHK(c, hipMalloc(&c->d_invj, sizeof(double) * E));
    HIPCHK(c, hipMalloc(&c->d_center, sizeof(double) * 2 * E));
    HIPCHK(c, hipMalloc(&c->d_length, sizeof(double) * E));
    HIPCHK(c, hipMalloc(&c->d_trace, sizeof(double) * (size_t)E * 4 * c->n * (c->sc ? kTraceQSC : kTraceQ)));
    HIPCHK(c, hipMalloc(&c->d_buf[0], sizeof(double) * nd));
    HIPCHK(c, hipMalloc(&c->d_buf[1], sizeof(double) * nd));
    HIPCHK(c, hipMalloc(&c->d_us, sizeof(double) * nd));
    // tables and partials sized for the smallest tile, so the tile can change freely
    c->vcap = (E + min_eb(c->N) - 1) / min_eb(c->N) + ninst;
    const int scap = (E + surf_eb(c->N) - 1) / surf_eb(c->N) + ninst;
    // hyperbolic-parabolic: the viscous partials follow in a second block of the same size
    c->pv_off = c->parabolic ? c->vcap + scap : 0;
    // partials, then per instance 2·kSlots CFL/viscous maxima slots, zero between reductions
    c->slot_off = (size_t)((c->parabolic ? 2 : 1) * (c->vcap + scap) + 1);
    const size_t npart = c->slot_off + (size_t)ninst * 2 * kSlots;
    HIPCHK(c, hipMalloc(&c->d_partial, sizeof(double) * npart));
    HIPCHK(c, hipMemset(c->d_partial, 0, sizeof(double) * npart));
    HIPCHK(c, hipMalloc(&c->d_vblk, sizeof(int4) * (size_t)c->vcap));
    HIPCHK(c, hipMalloc(&c->d_sblk, sizeof(int4) * (size_t)scap));
    HIPCHK(c, hipMalloc(&c->d_etile, sizeof(int) * (size_t)E));
    HIPCHK(c, hipMalloc(&c->d_rng, sizeof(int4) * (size_t)ninst));
    if (c->hconst.ns_walls) HIPCHK(c, hipMalloc(&c->d_fvb, sizeof(double) * (size_t)E * 4 * c->n * 6));
    if (pd) {
        HIPCHK(c, hipMalloc(&c->d_plink, sizeof(int4) * plink.size()));
        HIPCHK(c, hipMemcpy(c->d_plink, plink.data(), sizeof(int4) * plink.size(), hipMemcpyHostToDevice));
    }
    if (M > 0) {
        std::vector<int2> mo(M);
        for (int q = 0; q < M; ++q) mo[q] = make_int2(md->mortars_host[5 * q + 3], md->mortars_host[5 * q + 4]);
        HIPCHK(c, hipMalloc(&c->d_mort, sizeof(int2) * (size_t)M));
        HIPCHK(c, hipMalloc(&c->d_mrec, sizeof(double) * (size_t)M * 3 * c->n * 12));
        HIPCHK(c, hipMalloc(&c->d_mflux, sizeof(double) * (size_t)M * 3 * c->n * 4));
        HIPCHK(c, hipMemcpy(c->d_mort, mo.data(), sizeof(int2) * (size_t)M, hipMemcpyHostToDevice));
    }
    if (c->sc) {
        HIPCHK(c, hipMalloc(&c->d_alpha_tmp, sizeof(double) * (size_t)E));
        HIPCHK(c, hipMalloc(&c->d_alpha, sizeof(double) * (size_t)E));
    }
    if (c->flags & TLW_FLAG_KEEP_AUX) {
        HIPCHK(c, hipMalloc(&c->d_auxU, sizeof(double) * nd));
        HIPCHK(c, hipMalloc(&c->d_auxF, sizeof(double) * 2 * nd));
    }
    HIPCHK(c, hipMalloc(&c->d_state, sizeof(DevState) * ninst));
    HIPCHK(c, hipMalloc(&c->d_cur, sizeof(int) * ninst));
    HIPCHK(c, hipHostMalloc(&c->h_state, sizeof(DevState) * ninst, hipHostMallocDefault));
    std::memset(c->h_state, 0, sizeof(DevState) * ninst);
    for (int k = 0; k < ninst; ++k) {
        c->h_state[k].invJ0 = invj[c->inst_first[k]];
        c->h_state[k].ndofs = (double)ms[k].n_elements * c->n * c->n;  // Trixi.ndofs
    }
    HIPCHK(c, hipMemcpy(c->d_state, c->h_state, sizeof(DevState) * ninst, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_nbr, nbr.data(), sizeof(int) * 4 * E, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_invj, invj.data(), sizeof(double) * E, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_center, center.data(), sizeof(double) * 2 * E, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_length, length.data(), sizeof(double) * E, hipMemcpyHostToDevice));
    if (c->curved) {
        const size_t NN = (size_t)c->n * c->n;
        HIPCHK(c, hipMalloc(&c->d_Ja, sizeof(double) * 4 * NN * E));
        HIPCHK(c, hipMalloc(&c->d_ijn, sizeof(double) * NN * E));
        HIPCHK(c, hipMalloc(&c->d_fnorm, sizeof(double) * fnorm.size()));
        HIPCHK(c, hipMemcpy(c->d_Ja, cv->contravariant_vectors_host, sizeof(double) * 4 * NN * E,
                            hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->d_ijn, cv->inverse_jacobian_host, sizeof(double) * NN * E, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->d_fnorm, fnorm.data(), sizeof(double) * fnorm.size(), hipMemcpyHostToDevice));
        if (!fxy.empty()) {
            HIPCHK(c, hipMalloc(&c->d_fxy, sizeof(double) * fxy.size()));
            HIPCHK(c, hipMemcpy(c->d_fxy, fxy.data(), sizeof(double) * fxy.size(), hipMemcpyHostToDevice));
        }
        if (c->hconst.src_kind != TLW_SRC_NONE) {  // the cache's node_coordinates[m, i, j, e] as they are
            HIPCHK(c, hipMalloc(&c->d_nxy, sizeof(double) * 2 * NN * E));
            HIPCHK(c, hipMemcpy(c->d_nxy, cv->node_coordinates_host, sizeof(double) * 2 * NN * E,
                                hipMemcpyHostToDevice));
        }
        if (!snrm.empty()) {
            HIPCHK(c, hipMalloc(&c->d_snrm, sizeof(double) * snrm.size()));
            HIPCHK(c, hipMalloc(&c->d_sqrk, sizeof(double) * sqrk.size()));
            HIPCHK(c, hipMemcpy(c->d_snrm, snrm.data(), sizeof(double) * snrm.size(), hipMemcpyHostToDevice));
            HIPCHK(c, hipMemcpy(c->d_sqrk, sqrk.data(), sizeof(double) * sqrk.size(), hipMemcpyHostToDevice));
        }
    }
    if (c->hconst.bc_kind == TLW_BV_TABLE) {
        // boundary face nodes in table order: element, face (-x, +x, -y, +y), node
        const int n = c->n;
        std::vector<int> bidx(4 * (size_t)E, -1);
        long long nb = 0;
        for (size_t q = 0; q < 4 * (size_t)E; ++q)
            if (nbr[q] == -1) bidx[q] = (int)nb++;
        c->n_bnodes = nb * n;
        c->bxy.assign(2 * (size_t)c->n_bnodes, 0.0);
        for (int e = 0; e < E; ++e)
            for (int f = 0; f < 4; ++f) {
                const int b = bidx[4 * (size_t)e + f];
                if (b < 0) continue;
                for (int l = 0; l < n; ++l) {
                    double x, y;
                    if (c->curved) {
                        x = fxy[((((size_t)e * 4 + f) * n + l) * 2)];
                        y = fxy[((((size_t)e * 4 + f) * n + l) * 2) + 1];
                    } else {  // the device's TreeMesh face-node coordinates
                        const double half = 0.5 * length[e];
                        if (f < 2) {
                            x = center[2 * (size_t)e] + half * c->hconst.nodes[f == 0 ? 0 : n - 1];
                            y = center[2 * (size_t)e + 1] + half * c->hconst.nodes[l];
                        } else {
                            x = center[2 * (size_t)e] + half * c->hconst.nodes[l];
                            y = center[2 * (size_t)e + 1] + half * c->hconst.nodes[f == 2 ? 0 : n - 1];
                        }
                    }
                    c->bxy[2 * ((size_t)b * n + l)] = x;
                    c->bxy[2 * ((size_t)b * n + l) + 1] = y;
                }
            }
        HIPCHK(c, hipMalloc(&c->d_bidx, sizeof(int) * bidx.size()));
        HIPCHK(c, hipMemcpy(c->d_bidx, bidx.data(), sizeof(int) * bidx.size(), hipMemcpyHostToDevice));
    }
    return build_tables(c);
}

KArgs base_args(tlw_ctx* c) {
    KArgs a{};
    a.E = c->E;
    a.bufs[0] = c->d_buf[0];
    a.bufs[1] = c->d_buf[1];
    a.trace = c->d_trace;
    a.nbr = c->d_nbr;
    a.inv_jac = c->d_invj;
    a.center = c->d_center;
    a.length = c->d_length;
    a.c = c->d_const;
    a.st = c->d_state;
    a.partial = c->d_partial;
    a.pv_off = c->pv_off;
    a.vol_grid = c->vgrid;
    a.ebv = c->ebv;
    a.vblk = c->d_vblk;
    a.sblk = c->d_sblk;
    a.etile = c->d_etile;
    a.alpha_tmp = c->d_alpha_tmp;
    a.alpha_out = c->d_alpha;
    a.bc_scale = 1.0;
    a.Ja = c->d_Ja;
    a.ijn = c->d_ijn;
    a.fnorm = c->d_fnorm;
    a.fxy = c->d_fxy;
    a.nxy = c->d_nxy;
    a.slots = c->d_partial + c->slot_off;
    a.snrm = c->d_snrm;
    a.fvb = c->d_fvb;
    a.sqrk = c->d_sqrk;
    a.bidx = c->d_bidx;
    a.btab = c->d_btab;
    a.btab_levels = c->btab_levels;
    a.plink = c->d_plink;
    a.n_mortars = c->n_mortars;
    a.mort = c->d_mort;
    a.mrec = c->d_mrec;
    a.mflux = c->d_mflux;
    return a;
}

// Kernel launches, dispatched on the polynomial degree and the tile size.
template <int N, int EB, bool SC, bool PAR = false> void launch_volume_eb(tlw_ctx* c, const KArgs& a, int kind) {
    dim3 g(c->vgrid), b(VGeo<N, EB>::NT);
    if constexpr (!SC && !PAR) {  // the Euler kernels with source terms
        if (kind == VOL_LW_SRC) {
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_volume<N, EB, VOL_LW_SRC, false, false>), g, b, 0, c->stream, a);
            return;
        } else if (kind == VOL_MDRK1_SRC) {
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_volume<N, EB, VOL_MDRK1_SRC, false, false>), g, b, 0, c->stream, a);
            return;
        } else if (kind == VOL_MDRK2_SRC) {
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_volume<N, EB, VOL_MDRK2_SRC, false, false>), g, b, 0, c->stream, a);
            return;
        }
    }
    if (kind == VOL_LW)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_volume<N, EB, VOL_LW, SC, PAR>), g, b, 0, c->stream, a);
    else if (kind == VOL_MDRK1)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_volume<N, EB, VOL_MDRK1, SC, PAR>), g, b, 0, c->stream, a);
    else if (kind == VOL_MDRK2)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_volume<N, EB, VOL_MDRK2, SC, PAR>), g, b, 0, c->stream, a);
    else if (vol_is_wall(kind)) {
        // the LW wall kernel runs at one wave per SIMD: tiles of at most 4 waves
        if constexpr (PAR && !SC && VGeo<N, EB>::NT <= 256) {
            if (kind == VOL_LW_WALL)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_volume<N, EB, VOL_LW_WALL, SC, PAR>), g, b, 0, c->stream, a);
            else if (kind == VOL_MDRK1_WALL)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_volume<N, EB, VOL_MDRK1_WALL, SC, PAR>), g, b, 0, c->stream, a);
            else if (kind == VOL_MDRK2_WALL)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_volume<N, EB, VOL_MDRK2_WALL, SC, PAR>), g, b, 0, c->stream, a);
            else if constexpr (N == 4)  // VOL_LW_ADAPT_WALL: the temporal error is kernel 4's
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_volume<N, EB, VOL_LW_ADAPT_WALL, SC, PAR>), g, b, 0, c->stream,
                                   a);
            else
                c->bad_launch = true;
        } else {
            c->bad_launch = true;
        }
    } else if constexpr (N == 4)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_volume<N, EB, VOL_LW_ADAPT, SC, PAR>), g, b, 0, c->stream, a);
}
template <int N, bool SC, bool PAR = false> void launch_volume_sc(tlw_ctx* c, const KArgs& a, int kind) {
    constexpr const int* ch = N == 4 ? kEbChoices4 : kEbChoices3;
    if (c->ebv == ch[0]) launch_volume_eb<N, ch[0], SC, PAR>(c, a, kind);
    else if (c->ebv == ch[1]) launch_volume_eb<N, ch[1], SC, PAR>(c, a, kind);
    else launch_volume_eb<N, ch[2], SC, PAR>(c, a, kind);
}
template <int N, int EB, bool SC> void launch_volume_curved_eb(tlw_ctx* c, const KArgs& a, int kind) {
    dim3 g(c->vgrid), b(VGeo<N, EB>::NT);
    if constexpr (!SC) {  // the LW and MDRK kernels with source terms
        if (kind == VOL_LW_SRC) {
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_volume<N, EB, VOL_LW_SRC, false, false, true>), g, b, 0, c->stream,
                               a);
            return;
        } else if (kind == VOL_MDRK1_SRC) {
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_volume<N, EB, VOL_MDRK1_SRC, false, false, true>), g, b, 0,
                               c->stream, a);
            return;
        } else if (kind == VOL_MDRK2_SRC) {
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_volume<N, EB, VOL_MDRK2_SRC, false, false, true>), g, b, 0,
                               c->stream, a);
            return;
        }
    }
    if (kind == VOL_LW)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_volume<N, EB, VOL_LW, SC, false, true>), g, b, 0, c->stream, a);
    else if (kind == VOL_LW_ADAPT)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_volume<N, EB, VOL_LW_ADAPT, SC, false, true>), g, b, 0, c->stream, a);
    else if (kind == VOL_MDRK1)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_volume<N, EB, VOL_MDRK1, SC, false, true>), g, b, 0, c->stream, a);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_volume<N, EB, VOL_MDRK2, SC, false, true>), g, b, 0, c->stream, a);
}
template <int N, bool SC> void launch_volume_curved(tlw_ctx* c, const KArgs& a, int kind) {
    constexpr const int* ch = N == 4 ? kEbChoices4 : kEbChoices3;
    if (c->ebv == ch[0]) launch_volume_curved_eb<N, ch[0], SC>(c, a, kind);
    else if (c->ebv == ch[1]) launch_volume_curved_eb<N, ch[1], SC>(c, a, kind);
    else launch_volume_curved_eb<N, ch[2], SC>(c, a, kind);
}
// compressible Navier-Stokes on the curvilinear mesh
template <int N, int EB> void launch_volume_curved_par_eb(tlw_ctx* c, const KArgs& a, int kind) {
    dim3 g(c->vgrid), b(VGeo<N, EB>::NT);
    if (kind == VOL_LW)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_volume<N, EB, VOL_LW, false, true, true>), g, b, 0, c->stream, a);
    else if (kind == VOL_MDRK1)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_volume<N, EB, VOL_MDRK1, false, true, true>), g, b, 0, c->stream, a);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_volume<N, EB, VOL_MDRK2, false, true, true>), g, b, 0, c->stream, a);
}
template <int N> void launch_volume_curved_par(tlw_ctx* c, const KArgs& a, int kind) {
    constexpr const int* ch = N == 4 ? kEbChoices4 : kEbChoices3;
    if (c->ebv == ch[0]) launch_volume_curved_par_eb<N, ch[0]>(c, a, kind);
    else if (c->ebv == ch[1]) launch_volume_curved_par_eb<N, ch[1]>(c, a, kind);
    else launch_volume_curved_par_eb<N, ch[2]>(c, a, kind);
}
template <int N> void launch_volume_t(tlw_ctx* c, const KArgs& a, int kind) {
    if (c->curved) {  // LW or MDRK (adaptive steps are rejected on such a context)
        if (c->parabolic) {
            launch_volume_curved_par<N>(c, a, kind);
        } else if (c->sc) {
            hipLaunchKernelGGL(k_indicator<N>, dim3(c->sgrid), dim3(kSurfBlock), 0, c->stream, a, c->d_alpha_tmp);
            launch_volume_curved<N, true>(c, a, kind);
        } else {
            const bool src = c->hconst.src_kind != TLW_SRC_NONE && kind <= VOL_MDRK2;
            launch_volume_curved<N, false>(c, a, src ? kind + (int)VOL_LW_SRC : kind);
        }
    } else if (c->parabolic) {
        // walls: the instantiations that keep the wall face nodes' viscous averages
        const int wk = kind == VOL_LW         ? VOL_LW_WALL
                       : kind == VOL_MDRK1    ? VOL_MDRK1_WALL
                       : kind == VOL_MDRK2    ? VOL_MDRK2_WALL
                                              : VOL_LW_ADAPT_WALL;
        launch_volume_sc<N, false, true>(c, a, c->hconst.ns_walls ? wk : kind);
    } else if (c->sc) {
        // IndicatorHennemannGassner on uⁿ before the volume integral (dg_2d.jl:1665)
        hipLaunchKernelGGL(k_indicator<N>, dim3(c->sgrid), dim3(kSurfBlock), 0, c->stream, a, c->d_alpha_tmp);
        launch_volume_sc<N, true>(c, a, kind);
    } else if (c->hconst.src_kind != TLW_SRC_NONE && kind <= VOL_MDRK2) {
        launch_volume_sc<N, false>(c, a, kind + VOL_LW_SRC);  // LW, MDRK1, MDRK2 with source terms
    } else {
        launch_volume_sc<N, false>(c, a, kind);
    }
}
void launch_volume(tlw_ctx* c, const KArgs& a, int kind) {
    if (c->N == 4) launch_volume_t<4>(c, a, kind);
    else launch_volume_t<3>(c, a, kind);
}
FinArgs fin_args(tlw_ctx* c, int reduce, int use_vol) {
    return FinArgs{c->d_state, c->d_partial, c->d_rng, use_vol, reduce, c->pv_off, c->d_partial + c->slot_off,
                   DtLog{c->d_dtlog, c->dtlog_cap, c->d_dtpresc, c->dtpresc_n}};
}
void launch_finalize(tlw_ctx* c, int reduce, int use_vol) {
    const FinArgs f = fin_args(c, reduce, use_vol);
    if (c->N == 4)
        hipLaunchKernelGGL(k_finalize<4>, dim3(c->ninst), dim3(kFinalizeBlock), 0, c->stream, f);
    else
        hipLaunchKernelGGL(k_finalize<3>, dim3(c->ninst), dim3(kFinalizeBlock), 0, c->stream, f);
}
template <int N> void launch_surface_t(tlw_ctx* c, const KArgs& a) {
    dim3 g(c->sgrid), b(kSurfBlock);
    if (c->curved && c->has_boundary && a.in_sel == SEL_EXPLICIT && !a.u) {
        // the curvilinear boundary faces read uⁿ: never launch without it
        c->bad_launch = true;
        return;
    }
    if (c->n_mortars > 0) {  // mortar fluxes from the volume kernel's face-node values
        constexpr int MPB = kMortarBlock / (N + 1);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_mortar<N>), dim3((c->n_mortars + MPB - 1) / MPB), dim3(kMortarBlock), 0,
                           c->stream, a);
    }
    if (c->curved) {
        if (c->sc) {
            if (c->has_boundary) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_surface<N, true, true, true>), g, b, 0, c->stream, a);
            else hipLaunchKernelGGL(HIP_KERNEL_NAME(k_surface<N, false, true, true>), g, b, 0, c->stream, a);
        } else if (c->has_boundary) {
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_surface<N, true, false, true>), g, b, 0, c->stream, a);
        } else {
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_surface<N, false, false, true>), g, b, 0, c->stream, a);
        }
    } else if (c->n_mortars > 0) {  // plain FR on a TreeMesh with refined cells
        if (c->bc_faces)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_surface<N, true, false, false, true>), g, b, 0, c->stream, a);
        else
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_surface<N, false, false, false, true>), g, b, 0, c->stream, a);
    } else if (c->sc) {
        if (c->has_boundary) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_surface<N, true, true>), g, b, 0, c->stream, a);
        else hipLaunchKernelGGL(HIP_KERNEL_NAME(k_surface<N, false, true>), g, b, 0, c->stream, a);
    } else {
        if (c->has_boundary) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_surface<N, true, false>), g, b, 0, c->stream, a);
        else hipLaunchKernelGGL(HIP_KERNEL_NAME(k_surface<N, false, false>), g, b, 0, c->stream, a);
    }
    // the step's CFL partials come from the surface kernel alone; with adaptive
    // steps the volume kernel's temporal-error sums are read by the finalize kernel
    if (a.reduce != RED_NONE) launch_finalize(c, a.reduce == RED_ADAPT ? RED_STEP : a.reduce, 0);
}
void launch_surface(tlw_ctx* c, const KArgs& a) {
    if (c->N == 4) launch_surface_t<4>(c, a);
    else launch_surface_t<3>(c, a);
}
void launch_limiter(tlw_ctx* c, const KArgs& a) {
    if (c->N == 4) hipLaunchKernelGGL(k_limiter<4>, dim3(c->sgrid), dim3(kSurfBlock), 0, c->stream, a);
    else hipLaunchKernelGGL(k_limiter<3>, dim3(c->sgrid), dim3(kSurfBlock), 0, c->stream, a);
    if (a.reduce != RED_NONE) launch_finalize(c, a.reduce, 0);
}
void launch_max_speed(tlw_ctx* c, const KArgs& a) {
    if (c->N == 4) hipLaunchKernelGGL(k_max_speed<4>, dim3(c->sgrid), dim3(kSurfBlock), 0, c->stream, a);
    else hipLaunchKernelGGL(k_max_speed<3>, dim3(c->sgrid), dim3(kSurfBlock), 0, c->stream, a);
    launch_finalize(c, a.reduce, 0);
}

int check_launch(tlw_ctx* c) {
    if (c->bad_launch) {
        c->bad_launch = false;
        return fail(c, TLW_EINVAL, "internal error: surface kernel launched without its input state");
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(c, TLW_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return TLW_OK;
}

hipEvent_t next_event(tlw_ctx* c) {
    if (c->ev_used == c->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->ev_pool.push_back(e);
    }
    return c->ev_pool[c->ev_used++];
}

// Records an event pair around a launch when timing is on.
template <class F> void timed(tlw_ctx* c, int kind, F&& launch) {
    if (!c->timing) {
        launch();
        return;
    }
    const size_t i0 = c->ev_used;
    hipEvent_t a = next_event(c), b = next_event(c);
    if (!a || !b) {
        launch();
        return;
    }
    (void)hipEventRecord(a, c->stream);
    launch();
    (void)hipEventRecord(b, c->stream);
    c->ev_pairs.push_back({(int)i0, kind});
}

// Folds recorded event pairs into the totals (stream must be synchronised).
void collect_times(tlw_ctx* c) {
    for (auto& pr : c->ev_pairs) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->ev_pool[pr.first], c->ev_pool[pr.first + 1]) == hipSuccess) {
            c->t_ms[pr.second] += ms;
            c->t_n[pr.second] += 1;
        }
    }
    c->ev_pairs.clear();
    c->ev_used = 0;
}

// One rhs evaluation (volume + surface) on explicit pointers.
int rhs_common(tlw_ctx* c, int kind, const double* u, double* du, double t, double dt) {
    if (!c) return TLW_EINVAL;
    if (!u || !du) return fail(c, TLW_EINVAL, "null state pointer");
    if (!std::isfinite(dt) || !std::isfinite(t)) return fail(c, TLW_EINVAL, "t and dt must be finite");
    if (kind != VOL_LW && c->time_disc != TLW_MDRK)
        return fail(c, TLW_EINVAL, "rhs_mdrk needs a context created with TLW_MDRK");
    if (kind == VOL_LW && c->hconst.src_kind == TLW_SRC_NAVIER_STOKES_CONVERGENCE_TEST && !c->hconst.ns_walls)
        return fail(c, TLW_ENOTSUP, "the Navier-Stokes convergence source runs in the MDRK stages or with walls");
    if (c->hconst.bc_kind == TLW_BV_TABLE && c->has_boundary && c->btab_levels == 0)
        return fail(c, TLW_EINVAL, "no boundary data uploaded (tlw_set_boundary_data)");
    KArgs a = base_args(c);
    a.u = u;
    a.us = c->d_us;
    a.out = du;
    a.in_sel = SEL_EXPLICIT;
    a.out_sel = SEL_EXPLICIT;
    a.t_arg = t;
    a.dt_arg = dt;
    a.use_state_dt = 0;
    a.step_mode = 0;
    a.reduce = RED_NONE;
    // tree/multi_derivative_2d.jl:34; a StructuredMesh passes no scaling (structured/multi_derivative_2d.jl:18-20)
    a.bc_scale = (kind == VOL_MDRK1 && (!c->curved || c->p4est)) ? 0.5 : 1.0;
    if (c->flags & TLW_FLAG_KEEP_AUX) {
        a.auxU = c->d_auxU;
        a.auxF = c->d_auxF;
    }
    launch_volume(c, a, kind);
    launch_surface(c, a);
    int rc = check_launch(c);
    if (rc) return rc;
    c->aux_valid = (c->flags & TLW_FLAG_KEEP_AUX) != 0;
    c->alpha_valid = c->sc;
    return TLW_OK;
}

// The launches of one step of the device loop (solve_lwfr's perform_step! with
// compute_dt of the next step folded into the reductions).
void enqueue_step(tlw_ctx* c) {
    KArgs a = base_args(c);
    a.use_state_dt = 1;
    a.step_mode = 1;
    // Without a limiter the CFL reduction of uⁿ⁺¹ is split between the volume
    // kernel (nodes it finalises) and the surface kernel (the rest); with the
    // positivity limiter (applied after the update, lwfr.jl:226) the limiter
    // kernel takes it over the limited state.
    const int red = c->adaptive ? RED_ADAPT : (c->positivity ? RED_NONE : RED_STEP);
    if (c->time_disc == TLW_LW) {
        KArgs v = a;
        v.in_sel = SEL_CUR;
        v.out_sel = SEL_OTHER;
        v.reduce = red;
        timed(c, 0, [&] { launch_volume(c, v, c->adaptive ? VOL_LW_ADAPT : VOL_LW); });
        KArgs sf = a;
        sf.in_sel = SEL_CUR;  // uⁿ: read at curvilinear domain-boundary faces
        sf.out_sel = SEL_OTHER;
        sf.reduce = red;
        timed(c, 1, [&] { launch_surface(c, sf); });
    } else {
        // stage 1: us = uⁿ + dt rhs_mdrk1(uⁿ)  (mdrk.jl:13-17)
        KArgs v1 = a;
        v1.in_sel = SEL_CUR;
        v1.out_sel = SEL_EXPLICIT;
        v1.out = c->d_us;
        v1.reduce = RED_NONE;
        timed(c, 0, [&] { launch_volume(c, v1, VOL_MDRK1); });
        KArgs s1 = v1;
        s1.bc_scale = (c->curved && !c->p4est) ? 1.0 : 0.5;
        timed(c, 1, [&] { launch_surface(c, s1); });
        // stage 2: uⁿ⁺¹ = uⁿ + dt rhs_mdrk2(uⁿ, us)  (mdrk.jl:21-24)
        KArgs v2 = a;
        v2.in_sel = SEL_CUR;
        v2.out_sel = SEL_OTHER;
        v2.us = c->d_us;
        v2.reduce = red;
        timed(c, 0, [&] { launch_volume(c, v2, VOL_MDRK2); });
        KArgs s2 = v2;
        timed(c, 1, [&] { launch_surface(c, s2); });
    }
    if (c->positivity && !c->adaptive) {
        KArgs l = a;
        l.out_sel = SEL_OTHER;
        l.reduce = RED_STEP;
        timed(c, 2, [&] { launch_limiter(c, l); });
    }
}

// Replays max_steps steps from a graph captured on first use for this step
// count and stream.  The captured launches contain no memory sets.
int advance_graph(tlw_ctx* c, int max_steps) {
    if (!c->stream) return fail(c, TLW_EINVAL, "graph replay needs an explicit stream (tlw_set_stream)");
    if (c->graph_stream != c->stream) {
        drop_graphs(c);
        c->graph_stream = c->stream;
    }
    auto it = c->graph_cache.find(max_steps);
    if (it == c->graph_cache.end()) {
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            for (int k = 0; k < max_steps; ++k) enqueue_step(c);
            e = hipStreamEndCapture(c->stream, &g);
        }
        hipGraphExec_t ge = nullptr;
        if (e == hipSuccess) e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (e != hipSuccess) return fail(c, TLW_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
        it = c->graph_cache.emplace(max_steps, ge).first;
    }
    HIPCHK(c, hipGraphLaunch(it->second, c->stream));
    return TLW_OK;
}

int sync_state(tlw_ctx* c) {
    HIPCHK(c, hipMemcpyAsync(c->h_state, c->d_state, sizeof(DevState) * c->ninst, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TLW_OK;
}

bool any_bad(const tlw_ctx* c) {
    for (int k = 0; k < c->ninst; ++k)
        if (c->h_state[k].bad_state) return true;
    return false;
}

// The device loop computes dt itself, so caller-defined boundary data must hold for
// every time (one time level) there.
int check_loop_boundary(tlw_ctx* c) {
    if (c->hconst.bc_kind != TLW_BV_TABLE || !c->has_boundary) return TLW_OK;
    if (c->btab_levels == 0) return fail(c, TLW_EINVAL, "no boundary data uploaded (tlw_set_boundary_data)");
    if (c->btab_levels != 1)
        return fail(c, TLW_ENOTSUP, "time-dependent boundary data needs the host loop (the device loop sets dt itself)");
    return TLW_OK;
}

// KArgs of one LW step of the device loop (the arguments enqueue_step gives
// the volume and surface launches of a plain LW context).
KArgs lw_step_args(tlw_ctx* c) {
    KArgs a = base_args(c);
    a.use_state_dt = 1;
    a.step_mode = 1;
    a.in_sel = SEL_CUR;
    a.out_sel = SEL_OTHER;
    a.reduce = RED_STEP;
    return a;
}

// Lagrange interpolation matrix from nodes xin to points xout (barycentric
// form; a point on an input node takes that node's value).
void interp_matrix(const double* xin, int n, const double* xout, int m, double* out /* [m][n] */) {
    double wb[8];
    for (int k = 0; k < n; ++k) {
        wb[k] = 1.0;
        for (int q = 0; q < n; ++q)
            if (q != k) wb[k] *= xin[k] - xin[q];
        wb[k] = 1.0 / wb[k];
    }
    for (int a = 0; a < m; ++a) {
        int hit = -1;
        for (int k = 0; k < n; ++k)
            if (std::fabs(xout[a] - xin[k]) < 1e-15) hit = k;
        double den = 0.0;
        for (int k = 0; k < n; ++k) {
            out[a * n + k] = (hit >= 0) ? (k == hit ? 1.0 : 0.0) : wb[k] / (xout[a] - xin[k]);
            den += out[a * n + k];
        }
        if (hit < 0)
            for (int k = 0; k < n; ++k) out[a * n + k] /= den;
    }
}

// Trixi's LobattoLegendreMortarL2 for the basis nodes: forward_lower/upper
// interpolate a face polynomial to the GLL nodes of its lower/upper half;
// reverse_lower/upper project the halves' fluxes back onto the face in L2, by
// Gauss quadrature between GLL -> Gauss and Gauss -> GLL interpolations
// (Trixi.jl src/solvers/dgsem/l2projection.jl, calc_reverse_*(.., Val(:gauss))).
void mortar_operators(DevConst& h, int n) {
    double g[8], wg[8];
    for (int j = 0; j < n; ++j) {  // Gauss-Legendre nodes: Newton on P_n
        double x = -std::cos(M_PI * (j + 0.75) / (n + 0.5)), dP = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = x;
            for (int k = 2; k <= n; ++k) {
                const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dP = n * (x * p1 - p0) / (x * x - 1.0);
            const double d = p1 / dP;
            x -= d;
            if (std::fabs(d) < 1e-16) break;
        }
        g[j] = x;
        wg[j] = 2.0 / ((1.0 - x * x) * dP * dP);
    }
    double half[2][8], ghalf[2][8];
    for (int i = 0; i < n; ++i) {
        half[0][i] = 0.5 * (h.nodes[i] - 1.0);
        half[1][i] = 0.5 * (h.nodes[i] + 1.0);
        ghalf[0][i] = 0.5 * (g[i] - 1.0);
        ghalf[1][i] = 0.5 * (g[i] + 1.0);
    }
    double l2g[64], g2l[64], fw[64], lg[64];
    interp_matrix(h.nodes, n, g, n, l2g);
    interp_matrix(g, n, h.nodes, n, g2l);
    for (int s = 0; s < 2; ++s) {
        interp_matrix(h.nodes, n, half[s], n, fw);
        interp_matrix(g, n, ghalf[s], n, lg);  // lg[j][i] = ℓ_i(½(g_j ± 1))
        double op[64], t[64];
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) op[i * n + j] = 0.5 * lg[j * n + i] * wg[j] / wg[i];
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < n; ++k) {
                double acc = 0.0;
                for (int j = 0; j < n; ++j) acc += op[i * n + j] * l2g[j * n + k];
                t[i * n + k] = acc;
            }
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < n; ++k) {
                double acc = 0.0;
                for (int j = 0; j < n; ++j) acc += g2l[i * n + j] * t[j * n + k];
                h.mrev[s][i][k] = acc;
                h.mfwd[s][i][k] = fw[i * n + k];
            }
    }
}

// message of the last failed context-free call on this thread (tlw_last_error(NULL))
thread_local std::string g_noctx_err;

int create_common(int ninst, const tlw_mesh_desc* meshes, const tlw_basis_desc* basis, const tlw_solver_desc* solver,
                  tlw_ctx** out, const tlw_parabolic_desc* par = nullptr, const tlw_curved_desc* cv = nullptr,
                  const tlw_mortar_desc* md = nullptr, const tlw_p4est_desc* pd = nullptr);

}  // namespace

extern "C" {

int tlw_create(const tlw_mesh_desc* mesh, const tlw_basis_desc* basis, const tlw_solver_desc* solver,
               tlw_ctx** out) {
    return create_common(1, mesh, basis, solver, out);
}

int tlw_create_parabolic(const tlw_mesh_desc* mesh, const tlw_basis_desc* basis, const tlw_solver_desc* solver,
                         const tlw_parabolic_desc* parabolic, tlw_ctx** out) {
    if (!parabolic) {
        if (out) *out = nullptr;
        return TLW_EINVAL;
    }
    return create_common(1, mesh, basis, solver, out, parabolic);
}

int tlw_create_curved(const tlw_mesh_desc* mesh, const tlw_basis_desc* basis, const tlw_solver_desc* solver,
                      const tlw_curved_desc* metrics, tlw_ctx** out) {
    if (!metrics) {
        if (out) *out = nullptr;
        return TLW_EINVAL;
    }
    return create_common(1, mesh, basis, solver, out, nullptr, metrics);
}

int tlw_create_curved_parabolic(const tlw_mesh_desc* mesh, const tlw_basis_desc* basis,
                                const tlw_solver_desc* solver, const tlw_curved_desc* metrics,
                                const tlw_parabolic_desc* parabolic, tlw_ctx** out) {
    if (!metrics || !parabolic) {
        if (out) *out = nullptr;
        return TLW_EINVAL;
    }
    return create_common(1, mesh, basis, solver, out, parabolic, metrics);
}

int tlw_create_batch(int32_t n_instances, const tlw_mesh_desc* meshes, const tlw_basis_desc* basis,
                     const tlw_solver_desc* solver, tlw_ctx** out) {
    if (n_instances < 1 || !meshes) {
        if (out) *out = nullptr;
        return TLW_EINVAL;
    }
    return create_common(n_instances, meshes, basis, solver, out);
}

int tlw_n_instances(const tlw_ctx* c, int32_t* n) {
    if (!c || !n) return TLW_EINVAL;
    *n = c->ninst;
    return TLW_OK;
}

int tlw_rebuild_mesh(tlw_ctx* c, const tlw_mesh_desc* mesh) {
    if (!c) return TLW_EINVAL;
    if (c->ninst != 1) return fail(c, TLW_EINVAL, "tlw_rebuild_mesh needs a single-instance context");
    if (c->curved) return fail(c, TLW_ENOTSUP, "a curvilinear context keeps its mesh (create a new one)");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return upload_meshes(c, mesh, 1);
}

int tlw_create_p4est(const tlw_mesh_desc* mesh, const tlw_p4est_desc* p4est, const tlw_basis_desc* basis,
                     const tlw_solver_desc* solver, tlw_ctx** out) {
    if (!p4est) {
        if (out) *out = nullptr;
        return TLW_EINVAL;
    }
    const tlw_curved_desc metrics{p4est->contravariant_vectors_host, p4est->inverse_jacobian_host,
                                  p4est->node_coordinates_host};
    return create_common(1, mesh, basis, solver, out, nullptr, &metrics, nullptr, p4est);
}

int tlw_create_nonconforming(const tlw_mesh_desc* mesh, const tlw_mortar_desc* mortars, const tlw_basis_desc* basis,
                             const tlw_solver_desc* solver, tlw_ctx** out) {
    if (!mortars) {
        if (out) *out = nullptr;
        return TLW_EINVAL;
    }
    return create_common(1, mesh, basis, solver, out, nullptr, nullptr, mortars);
}

int tlw_transfer_solution(tlw_ctx* c, const double* u_old_dev, int32_t n_old, double* u_new_dev, int32_t n_new,
                          const int32_t* plan_host) {
    if (!c || !u_old_dev || !u_new_dev || !plan_host || n_old <= 0 || n_new <= 0) return TLW_EINVAL;
    if (u_old_dev == u_new_dev) return fail(c, TLW_EINVAL, "tlw_transfer_solution: the states must not alias");
    for (int e = 0; e < n_new; ++e) {
        const int32_t* p = plan_host + 5 * (size_t)e;
        const bool ok = (p[0] == 0 && p[1] >= 0 && p[1] < n_old) ||
                        (p[0] == 1 && p[1] >= 0 && p[1] < n_old && p[2] >= 0 && p[2] < 4) ||
                        (p[0] == 2 && p[1] >= 0 && p[1] < n_old && p[2] >= 0 && p[2] < n_old && p[3] >= 0 &&
                         p[3] < n_old && p[4] >= 0 && p[4] < n_old);
        if (!ok) return fail(c, TLW_EINVAL, "tlw_transfer_solution: plan entry out of range");
    }
    int* d_plan = nullptr;
    HIPCHK(c, hipMalloc(&d_plan, sizeof(int32_t) * 5 * (size_t)n_new));
    hipError_t e1 = hipMemcpyAsync(d_plan, plan_host, sizeof(int32_t) * 5 * (size_t)n_new, hipMemcpyHostToDevice,
                                   c->stream);
    if (e1 == hipSuccess) {
        const long long total = (long long)n_new * c->n * c->n;
        const dim3 g((unsigned)((total + kTransferBlock - 1) / kTransferBlock)), b(kTransferBlock);
        if (c->N == 4) hipLaunchKernelGGL(k_transfer<4>, g, b, 0, c->stream, u_old_dev, u_new_dev, d_plan, n_new, c->d_const);
        else hipLaunchKernelGGL(k_transfer<3>, g, b, 0, c->stream, u_old_dev, u_new_dev, d_plan, n_new, c->d_const);
        e1 = hipGetLastError();
    }
    const hipError_t e2 = hipStreamSynchronize(c->stream);
    (void)hipFree(d_plan);
    if (e1 != hipSuccess) return fail(c, TLW_EHIP, std::string("tlw_transfer_solution: ") + hipGetErrorString(e1));
    if (e2 != hipSuccess) return fail(c, TLW_EHIP, std::string("tlw_transfer_solution: ") + hipGetErrorString(e2));
    return TLW_OK;
}

int tlw_rebuild_mesh_nonconforming(tlw_ctx* c, const tlw_mesh_desc* mesh, const tlw_mortar_desc* mortars) {
    if (!c || !mortars) return TLW_EINVAL;
    if (c->ninst != 1) return fail(c, TLW_EINVAL, "tlw_rebuild_mesh_nonconforming needs a single-instance context");
    if (c->curved) return fail(c, TLW_ENOTSUP, "a curvilinear context keeps its mesh (create a new one)");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return upload_meshes(c, mesh, 1, nullptr, mortars);
}

void tlw_destroy(tlw_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    else (void)hipDeviceSynchronize();
    free_mesh(c);
    for (auto e : c->ev_pool) (void)hipEventDestroy(e);
    (void)hipFree(c->d_dtlog);
    (void)hipFree(c->d_dtpresc);
    (void)hipFree(c->d_adj_tmp);
    (void)hipFree(c->d_adj_trace);
    (void)hipFree(c->d_const);
    delete c;
}

const char* tlw_last_error(const tlw_ctx* c) {
    if (c) return c->err.c_str();
    return g_noctx_err.empty() ? "null context" : g_noctx_err.c_str();
}

int tlw_set_stream(tlw_ctx* c, void* stream) {
    if (!c) return TLW_EINVAL;
    c->stream = static_cast<hipStream_t>(stream);
    return TLW_OK;
}

size_t tlw_state_len(const tlw_ctx* c) { return c ? state_len(c) : 0; }

int tlw_rhs_lw(tlw_ctx* c, const double* u, double* du, double t, double dt) {
    return rhs_common(c, VOL_LW, u, du, t, dt);
}
int tlw_rhs_mdrk1(tlw_ctx* c, const double* u, double* du, double t, double dt) {
    return rhs_common(c, VOL_MDRK1, u, du, t, dt);
}
int tlw_rhs_mdrk2(tlw_ctx* c, const double* u, double* du, double t, double dt) {
    return rhs_common(c, VOL_MDRK2, u, du, t, dt);
}

int tlw_us_ptr(tlw_ctx* c, double** us) {
    if (!c || !us) return TLW_EINVAL;
    *us = c->d_us;
    return TLW_OK;
}

int tlw_set_us(tlw_ctx* c, const double* us) {
    if (!c || !us) return fail(c, TLW_EINVAL, "null pointer");
    HIPCHK(c, hipMemcpyAsync(c->d_us, us, sizeof(double) * state_len(c), hipMemcpyDeviceToDevice, c->stream));
    return TLW_OK;
}

int tlw_max_dt(tlw_ctx* c, const double* u, double* max_dt) {
    if (!c || !u || !max_dt) return fail(c, TLW_EINVAL, "null pointer");
    hipLaunchKernelGGL(k_clear_bad_state, dim3((c->ninst + 255) / 256), dim3(256), 0, c->stream, c->d_state,
                       c->ninst);
    KArgs a = base_args(c);
    a.u = u;
    a.in_sel = SEL_EXPLICIT;
    a.reduce = RED_QUERY;
    launch_max_speed(c, a);
    int rc = check_launch(c);
    if (rc) return rc;
    rc = sync_state(c);
    if (rc) return rc;
    // a batch answers with the smallest of its instances' values
    double m = c->h_state[0].query;
    for (int k = 1; k < c->ninst; ++k) m = std::min(m, c->h_state[k].query);
    *max_dt = m;
    if (any_bad(c)) return fail(c, TLW_EADMISSIBLE, "non-admissible state (rho <= 0 or p <= 0)");
    return TLW_OK;
}

int tlw_set_state(tlw_ctx* c, const double* u, size_t len) {
    if (!c || !u) return fail(c, TLW_EINVAL, "null pointer");
    if (len != state_len(c)) return fail(c, TLW_EINVAL, "state length mismatch");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < c->ninst; ++k)
        HIPCHK(c, hipMemcpy(c->d_buf[c->host_cur[k]] + inst_offset(c, k), u + inst_offset(c, k),
                            sizeof(double) * inst_len(c, k), hipMemcpyHostToDevice));
    return TLW_OK;
}

int tlw_get_state(tlw_ctx* c, double* u, size_t len) {
    if (!c || !u) return fail(c, TLW_EINVAL, "null pointer");
    if (len != state_len(c)) return fail(c, TLW_EINVAL, "state length mismatch");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < c->ninst; ++k)
        HIPCHK(c, hipMemcpy(u + inst_offset(c, k), c->d_buf[c->host_cur[k]] + inst_offset(c, k),
                            sizeof(double) * inst_len(c, k), hipMemcpyDeviceToHost));
    return TLW_OK;
}

int tlw_set_state_dev(tlw_ctx* c, const double* u) {
    if (!c || !u) return fail(c, TLW_EINVAL, "null pointer");
    for (int k = 0; k < c->ninst; ++k)
        HIPCHK(c, hipMemcpyAsync(c->d_buf[c->host_cur[k]] + inst_offset(c, k), u + inst_offset(c, k),
                                 sizeof(double) * inst_len(c, k), hipMemcpyDeviceToDevice, c->stream));
    return TLW_OK;
}

int tlw_get_state_dev(tlw_ctx* c, double* u) {
    if (!c || !u) return fail(c, TLW_EINVAL, "null pointer");
    for (int k = 0; k < c->ninst; ++k)
        HIPCHK(c, hipMemcpyAsync(u + inst_offset(c, k), c->d_buf[c->host_cur[k]] + inst_offset(c, k),
                                 sizeof(double) * inst_len(c, k), hipMemcpyDeviceToDevice, c->stream));
    return TLW_OK;
}

int tlw_state_ptr(tlw_ctx* c, double** u) {
    if (!c || !u) return TLW_EINVAL;
    for (int k = 1; k < c->ninst; ++k)
        if (c->host_cur[k] != c->host_cur[0])
            return fail(c, TLW_EINVAL, "the instances' states are in different buffers (use tlw_get_state)");
    *u = c->d_buf[c->host_cur[0]];
    return TLW_OK;
}

int tlw_solve_init(tlw_ctx* c, double t0, double t_final, double cfl) {
    if (!c) return TLW_EINVAL;
    if (!std::isfinite(t0) || std::isnan(t_final) || !(cfl > 0.0))
        return fail(c, TLW_EINVAL, "t0 finite, t_final not NaN and cfl > 0 required");
    if (int rcb = check_loop_boundary(c)) return rcb;
    HIPCHK(c, hipMemcpyAsync(c->d_cur, c->host_cur.data(), sizeof(int) * c->ninst, hipMemcpyHostToDevice,
                             c->stream));
    if (c->adaptive) drop_graphs(c);
    c->adaptive = false;
    hipLaunchKernelGGL(k_init_state, dim3((c->ninst + 255) / 256), dim3(256), 0, c->stream, c->d_state, c->ninst, t0,
                       t_final, cfl, c->d_cur, 0, 0.0, 0.0, 0.0);
    KArgs a = base_args(c);
    a.in_sel = SEL_CUR;
    a.reduce = RED_INIT;
    launch_max_speed(c, a);
    int rc = check_launch(c);
    if (rc) return rc;
    rc = sync_state(c);
    if (rc) return rc;
    if (any_bad(c)) return fail(c, TLW_EADMISSIBLE, "non-admissible initial state");
    return TLW_OK;
}

int tlw_solve_init_adaptive(tlw_ctx* c, double t0, double t_final, double dt_initial, double abstol,
                            double reltol) {
    if (!c) return TLW_EINVAL;
    if (!std::isfinite(t0) || std::isnan(t_final) || !(dt_initial > 0.0) || !std::isfinite(dt_initial))
        return fail(c, TLW_EINVAL, "t0 finite, t_final not NaN and a finite dt_initial > 0 required");
    if (!(abstol >= 0.0) || !(reltol >= 0.0) || !(abstol > 0.0 || reltol > 0.0) || !std::isfinite(abstol) ||
        !std::isfinite(reltol))
        return fail(c, TLW_EINVAL, "tolerances: abstol, reltol >= 0, not both zero");
    if (c->time_disc != TLW_LW || (c->N != 4 && !c->curved) || (c->parabolic && c->curved))
        return fail(c, TLW_ENOTSUP,
                    "adaptive steps need the temporal-error estimate of an LW kernel (Euler on a TreeMesh at polydeg "
                    "4 or on a curvilinear mesh at polydeg 3 or 4; Navier-Stokes on a TreeMesh at polydeg 4, periodic "
                    "or with walls)");
    if (c->positivity) return fail(c, TLW_ENOTSUP, "adaptive steps are not combined with the positivity limiter");
    if (c->hconst.src_kind != TLW_SRC_NONE)
        return fail(c, TLW_ENOTSUP, "adaptive steps are not combined with source terms");
    if (int rcb = check_loop_boundary(c)) return rcb;
    HIPCHK(c, hipMemcpyAsync(c->d_cur, c->host_cur.data(), sizeof(int) * c->ninst, hipMemcpyHostToDevice,
                             c->stream));
    if (!c->adaptive) drop_graphs(c);
    c->adaptive = true;
    hipLaunchKernelGGL(k_init_state, dim3((c->ninst + 255) / 256), dim3(256), 0, c->stream, c->d_state, c->ninst, t0,
                       t_final, 1.0, c->d_cur, 1, dt_initial, abstol, reltol);
    int rc = check_launch(c);
    if (rc) return rc;
    return sync_state(c);
}

int tlw_rejected_steps(tlw_ctx* c, int64_t* nreject) {
    if (!c || !nreject) return fail(c, TLW_EINVAL, "null pointer");
    int rc = sync_state(c);
    if (rc) return rc;
    int64_t r = 0;
    for (int k = 0; k < c->ninst; ++k) r += c->h_state[k].nreject;
    *nreject = r;
    return TLW_OK;
}

int tlw_advance(tlw_ctx* c, int32_t max_steps, double* t_host, double* dt_host, int64_t* iter_host,
                int32_t* finished_host) {
    if (!c) return TLW_EINVAL;
    if (max_steps < 0) return fail(c, TLW_EINVAL, "max_steps < 0");
    if (max_steps > 0) {
        if (int rcb = check_loop_boundary(c)) return rcb;
        if (c->graphs && !c->timing) {
            int rc = advance_graph(c, max_steps);
            if (rc) return rc;
        } else {
            for (int s = 0; s < max_steps; ++s) enqueue_step(c);
        }
    }
    int rc = check_launch(c);
    if (rc) return rc;
    rc = sync_state(c);
    if (rc) return rc;
    if (c->timing) collect_times(c);
    for (int k = 0; k < c->ninst; ++k) c->host_cur[k] = c->h_state[k].cur;
    const DevState& s = c->h_state[0];
    if (t_host) *t_host = s.t;
    if (dt_host) *dt_host = s.dt;
    if (iter_host) *iter_host = s.iter;
    if (finished_host) {
        int all = 1;
        for (int k = 0; k < c->ninst; ++k) all = all && c->h_state[k].finished;
        *finished_host = all;
    }
    if (max_steps > 0) c->alpha_valid = c->sc;
    if (any_bad(c)) return fail(c, TLW_EADMISSIBLE, "non-admissible state (rho <= 0 or p <= 0)");
    return TLW_OK;
}

int tlw_batch_status(tlw_ctx* c, int32_t n, double* t, double* dt, int64_t* iter, int32_t* finished) {
    if (!c) return TLW_EINVAL;
    if (n != c->ninst) return fail(c, TLW_EINVAL, "n must equal the number of instances");
    int rc = sync_state(c);
    if (rc) return rc;
    for (int k = 0; k < n; ++k) {
        if (t) t[k] = c->h_state[k].t;
        if (dt) dt[k] = c->h_state[k].dt;
        if (iter) iter[k] = c->h_state[k].iter;
        if (finished) finished[k] = c->h_state[k].finished;
    }
    return TLW_OK;
}

int tlw_enable_graphs(tlw_ctx* c, int32_t enable) {
    if (!c) return TLW_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->graphs = enable != 0;
    if (!c->graphs) drop_graphs(c);
    return TLW_OK;
}

int tlw_enable_kernel_timing(tlw_ctx* c, int32_t enable) {
    if (!c) return TLW_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    collect_times(c);
    c->timing = enable != 0;
    return TLW_OK;
}

int tlw_kernel_times(tlw_ctx* c, double* vol_ms, int64_t* vol_n, double* surf_ms, int64_t* surf_n) {
    if (!c) return TLW_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    collect_times(c);
    if (vol_ms) *vol_ms = c->t_ms[0];
    if (vol_n) *vol_n = c->t_n[0];
    if (surf_ms) *surf_ms = c->t_ms[1];
    if (surf_n) *surf_n = c->t_n[1];
    c->t_ms[0] = c->t_ms[1] = c->t_ms[2] = 0.0;
    c->t_n[0] = c->t_n[1] = c->t_n[2] = 0;
    return TLW_OK;
}

int tlw_adjoint_lw(tlw_ctx* c, const double* u, const double* lam_next, double* lam, double t, double dt) {
    if (!c) return TLW_EINVAL;
    if (!u || !lam_next || !lam) return fail(c, TLW_EINVAL, "null state pointer");
    if (!std::isfinite(dt) || !std::isfinite(t)) return fail(c, TLW_EINVAL, "t and dt must be finite");
    if (c->time_disc != TLW_LW || c->parabolic || c->curved || c->has_boundary ||
        c->hconst.flux_order != TLW_FLUX_TREE || c->ninst != 1 || (c->sc && c->hconst.recon != TLW_RECON_FIRST_ORDER) ||
        c->hconst.src_kind != TLW_SRC_NONE)
        return fail(c, TLW_ENOTSUP,
                    "the adjoint covers LW (plain or blended with first-order subcell FV) on one periodic mesh with "
                    "the TreeMesh flux order and no source terms");
    if (lam == lam_next || lam == u) return fail(c, TLW_EINVAL, "lam must not alias lam_next or u");
    if (c->sc) {
        // blending factors of the step's state (frozen), then the blended step's adjoint
        if (!c->d_adj_trace)
            HIPCHK(c, hipMalloc(&c->d_adj_trace, sizeof(double) * (size_t)c->E * 4 * c->n * kAdjQSC));
        KArgs ia = base_args(c);
        ia.u = u;
        ia.in_sel = SEL_EXPLICIT;
        ia.use_state_dt = 0;
        AdjArgs aa{c->E, u, nullptr, lam_next, nullptr, lam, nullptr, c->d_adj_trace, c->d_nbr, c->d_invj, c->d_const,
                   dt};
        const double J0 = c->h_state[0].invJ0;
        if (c->N == 4) {
            constexpr int EB = kEBA<4>;
            dim3 g((c->E + EB - 1) / EB), b(VGeo<4, EB>::NT);
            hipLaunchKernelGGL(k_indicator<4>, dim3(c->sgrid), dim3(kSurfBlock), 0, c->stream, ia, c->d_alpha_tmp);
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_adj_records_sc<4, EB>), g, b, 0, c->stream, aa, c->d_alpha_tmp);
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_adjoint_sc<4, EB>), g, b, 0, c->stream, aa, c->d_alpha_tmp, J0);
        } else {
            constexpr int EB = kEBA<3>;
            dim3 g((c->E + EB - 1) / EB), b(VGeo<3, EB>::NT);
            hipLaunchKernelGGL(k_indicator<3>, dim3(c->sgrid), dim3(kSurfBlock), 0, c->stream, ia, c->d_alpha_tmp);
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_adj_records_sc<3, EB>), g, b, 0, c->stream, aa, c->d_alpha_tmp);
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_adjoint_sc<3, EB>), g, b, 0, c->stream, aa, c->d_alpha_tmp, J0);
        }
        return check_launch(c);
    }
    KArgs a = base_args(c);
    a.u = u;
    a.us = lam_next;
    a.out = lam;
    a.t_arg = t;
    a.dt_arg = dt;
    if (c->N == 4) {
        constexpr int EB = kEBA<4>;
        dim3 g((c->E + EB - 1) / EB), b(VGeo<4, EB>::NT);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_adj_records<4, EB>), g, b, 0, c->stream, a);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_adjoint<4, EB>), g, b, 0, c->stream, a);
    } else {
        constexpr int EB = kEBA<3>;
        dim3 g((c->E + EB - 1) / EB), b(VGeo<3, EB>::NT);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_adj_records<3, EB>), g, b, 0, c->stream, a);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_adjoint<3, EB>), g, b, 0, c->stream, a);
    }
    return check_launch(c);
}

int tlw_adjoint_mdrk(tlw_ctx* c, const double* u, const double* us, const double* lam_next, double* lam, double t,
                     double dt) {
    if (!c) return TLW_EINVAL;
    if (!u || !us || !lam_next || !lam) return fail(c, TLW_EINVAL, "null state pointer");
    if (!std::isfinite(dt) || !std::isfinite(t)) return fail(c, TLW_EINVAL, "t and dt must be finite");
    if (c->time_disc != TLW_MDRK || c->sc || c->parabolic || c->curved || c->has_boundary ||
        c->hconst.flux_order != TLW_FLUX_TREE || c->ninst != 1 || c->hconst.src_kind != TLW_SRC_NONE)
        return fail(c, TLW_ENOTSUP,
                    "the MDRK adjoint covers plain MDRK on one periodic mesh with the TreeMesh flux order and no "
                    "source terms");
    if (lam == lam_next || lam == u || lam == us) return fail(c, TLW_EINVAL, "lam must not alias u, us or lam_next");
    if (!c->d_adj_tmp) HIPCHK(c, hipMalloc(&c->d_adj_tmp, sizeof(double) * state_len(c)));
    AdjArgs a2{c->E, u, us, lam_next, nullptr, lam, c->d_adj_tmp, c->d_trace, c->d_nbr, c->d_invj, c->d_const, dt};
    AdjArgs a1{c->E, u, nullptr, c->d_adj_tmp, lam, lam, nullptr, c->d_trace, c->d_nbr, c->d_invj, c->d_const, dt};
    if (c->N == 4) {
        constexpr int EB = kEBA<4>;
        dim3 g((c->E + EB - 1) / EB), b(VGeo<4, EB>::NT);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_adj_records_md<4, EB, 2>), g, b, 0, c->stream, a2);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_adjoint_md<4, EB, 2>), g, b, 0, c->stream, a2);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_adj_records_md<4, EB, 1>), g, b, 0, c->stream, a1);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_adjoint_md<4, EB, 1>), g, b, 0, c->stream, a1);
    } else {
        constexpr int EB = kEBA<3>;
        dim3 g((c->E + EB - 1) / EB), b(VGeo<3, EB>::NT);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_adj_records_md<3, EB, 2>), g, b, 0, c->stream, a2);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_adjoint_md<3, EB, 2>), g, b, 0, c->stream, a2);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_adj_records_md<3, EB, 1>), g, b, 0, c->stream, a1);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_adjoint_md<3, EB, 1>), g, b, 0, c->stream, a1);
    }
    return check_launch(c);
}

int tlw_set_positivity_limiter(tlw_ctx* c, int32_t enable, double threshold_density, double threshold_pressure) {
    if (!c) return TLW_EINVAL;
    if (enable && (std::isnan(threshold_density) || std::isnan(threshold_pressure)))
        return fail(c, TLW_EINVAL, "positivity thresholds must not be NaN");
    if (enable && c->adaptive)
        return fail(c, TLW_ENOTSUP, "the positivity limiter is not combined with adaptive steps (tlw_solve_init first)");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if ((enable != 0) != c->positivity) drop_graphs(c);
    c->positivity = enable != 0;
    if (c->positivity) {
        c->hconst.pos_threshold[0] = threshold_density;
        c->hconst.pos_threshold[1] = threshold_pressure;
        HIPCHK(c, hipMemcpy(c->d_const, &c->hconst, sizeof(DevConst), hipMemcpyHostToDevice));
    }
    return TLW_OK;
}

int tlw_apply_positivity_limiter(tlw_ctx* c, double* u) {
    if (!c || !u) return fail(c, TLW_EINVAL, "null pointer");
    if (!c->positivity) return fail(c, TLW_EINVAL, "positivity limiter not configured (tlw_set_positivity_limiter)");
    KArgs a = base_args(c);
    a.out = u;
    a.out_sel = SEL_EXPLICIT;
    a.reduce = RED_NONE;
    launch_limiter(c, a);
    return check_launch(c);
}

int tlw_set_elements_per_workgroup(tlw_ctx* c, int32_t elements) {
    if (!c) return TLW_EINVAL;
    if (!eb_supported(c->N, elements))
        return fail(c, TLW_EINVAL, c->N == 4 ? "elements per workgroup must be 5, 10 or 15 for polydeg 4"
                                             : "elements per workgroup must be 4, 8 or 16 for polydeg 3");
    if (c->hconst.ns_walls && elements * c->n * c->n > 256)
        return fail(c, TLW_EINVAL, "Navier-Stokes walls: at most 256 nodes per workgroup (polydeg 4: 5 or 10)");
    if (elements == c->ebv) return TLW_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->ebv = elements;
    return build_tables(c);
}

int tlw_get_elements_per_workgroup(const tlw_ctx* c, int32_t* elements) {
    if (!c || !elements) return TLW_EINVAL;
    *elements = c->ebv;
    return TLW_OK;
}

int tlw_autotune(tlw_ctx* c, int32_t reps, int32_t* chosen, double* ms_out) {
    if (!c) return TLW_EINVAL;
    if (reps < 1) return fail(c, TLW_EINVAL, "reps >= 1 required");
    for (int k = 1; k < c->ninst; ++k)
        if (c->host_cur[k] != c->host_cur[0])
            return fail(c, TLW_EINVAL, "the instances' states are in different buffers");
    // one rhs evaluation of the context's scheme (LW, or MDRK stage 2, the
    // heavier stage) on the current state buffer, into the other buffer
    const int kind = c->time_disc == TLW_LW ? VOL_LW : VOL_MDRK2;
    const double* u = c->d_buf[c->host_cur[0]];
    double* scratch = c->d_buf[1 - c->host_cur[0]];
    if (kind == VOL_MDRK2)
        HIPCHK(c, hipMemcpyAsync(c->d_us, u, sizeof(double) * state_len(c), hipMemcpyDeviceToDevice, c->stream));
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(c, TLW_EHIP, "hipEventCreate failed");
    }
    const int saved = c->ebv;
    const double dt = 1e-6;  // any finite step: the work does not depend on it
    int best = saved;
    double best_ms = 0.0;
    int rc = TLW_OK;
    for (int k = 0; k < 3 && rc == TLW_OK; ++k) {
        if (c->hconst.ns_walls && eb_choices(c->N)[k] * c->n * c->n > 256) continue;  // not a wall tile
        rc = tlw_set_elements_per_workgroup(c, eb_choices(c->N)[k]);
        if (rc) break;
        rc = rhs_common(c, kind, u, scratch, 0.0, dt);  // warm-up
        if (rc) break;
        if (hipEventRecord(e0, c->stream) != hipSuccess) {
            rc = fail(c, TLW_EHIP, "hipEventRecord");
            break;
        }
        for (int r = 0; r < reps && rc == TLW_OK; ++r) rc = rhs_common(c, kind, u, scratch, 0.0, dt);
        if (rc) break;
        float ms = 0.f;
        if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            rc = fail(c, TLW_EHIP, "event timing failed");
            break;
        }
        const double per = ms / reps;
        if (ms_out) ms_out[k] = per;
        if (k == 0 || per < best_ms) {
            best_ms = per;
            best = c->ebv;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const int rc2 = tlw_set_elements_per_workgroup(c, rc == TLW_OK ? best : saved);
    c->aux_valid = false;
    c->alpha_valid = false;
    if (chosen) *chosen = c->ebv;
    return rc ? rc : rc2;
}

int tlw_get_aux(tlw_ctx* c, int32_t which, double* out, size_t len) {
    if (!c || !out) return fail(c, TLW_EINVAL, "null pointer");
    const size_t nd = state_len(c);
    const double* src = nullptr;
    size_t need = 0;
    if (which == TLW_AUX_ALPHA) {
        if (!c->sc) return fail(c, TLW_EINVAL, "context created without shock capturing");
        if (!c->alpha_valid) return fail(c, TLW_EINVAL, "no rhs evaluated or step taken yet");
        src = c->d_alpha;
        need = (size_t)c->E;
    } else if (which == TLW_AUX_U || which == TLW_AUX_F) {
        if (!(c->flags & TLW_FLAG_KEEP_AUX))
            return fail(c, TLW_EINVAL, "context created without TLW_FLAG_KEEP_AUX");
        if (!c->aux_valid) return fail(c, TLW_EINVAL, "no rhs evaluated yet");
        src = which == TLW_AUX_U ? c->d_auxU : c->d_auxF;
        need = which == TLW_AUX_U ? nd : 2 * nd;
    } else {
        return fail(c, TLW_EINVAL, "unknown aux selector");
    }
    if (len != need) return fail(c, TLW_EINVAL, "aux length mismatch");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(out, src, sizeof(double) * len, hipMemcpyDeviceToHost));
    return TLW_OK;
}

int tlw_boundary_nodes(tlw_ctx* c, int64_t* n_nodes, double* xy_host, size_t len) {
    if (!c || !n_nodes) return fail(c, TLW_EINVAL, "null pointer");
    if (c->hconst.bc_kind != TLW_BV_TABLE)
        return fail(c, TLW_EINVAL, "context created without caller-defined boundary data (TLW_BV_TABLE)");
    *n_nodes = c->n_bnodes;
    if (!xy_host) return TLW_OK;
    if (len != c->bxy.size()) return fail(c, TLW_EINVAL, "coordinate array length must be 2 * n_nodes");
    std::memcpy(xy_host, c->bxy.data(), sizeof(double) * len);
    return TLW_OK;
}

int tlw_set_boundary_data(tlw_ctx* c, const double* values, size_t len, int32_t time_levels) {
    if (!c || !values) return fail(c, TLW_EINVAL, "null pointer");
    if (c->hconst.bc_kind != TLW_BV_TABLE)
        return fail(c, TLW_EINVAL, "context created without caller-defined boundary data (TLW_BV_TABLE)");
    if (time_levels != 1 && time_levels != c->n + 1)
        return fail(c, TLW_EINVAL, "time_levels must be 1 (steady) or polydeg + 2 (GLL time levels, then t)");
    const size_t need = (size_t)c->n_bnodes * time_levels * 4;
    if (len != need) return fail(c, TLW_EINVAL, "boundary data length must be 4 * time_levels * n_nodes");
    for (size_t k = 0; k < len; k += 4)
        if (!(values[k] > 0.0) || !std::isfinite(values[k]))
            return fail(c, TLW_EINVAL, "boundary data: density must be finite and > 0");
    if (need > c->btab_cap) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        drop_graphs(c);
        (void)hipFree(c->d_btab);
        c->d_btab = nullptr;
        c->btab_cap = 0;
        HIPCHK(c, hipMalloc(&c->d_btab, sizeof(double) * (need > 0 ? need : 4)));
        c->btab_cap = need;
    }
    if (time_levels != c->btab_levels) drop_graphs(c);
    if (need > 0)
        HIPCHK(c, hipMemcpyAsync(c->d_btab, values, sizeof(double) * need, hipMemcpyHostToDevice, c->stream));
    c->btab_levels = time_levels;
    return TLW_OK;
}

int tlw_record_dts(tlw_ctx* c, int64_t capacity) {
    if (!c) return TLW_EINVAL;
    if (capacity < 0) return fail(c, TLW_EINVAL, "capacity < 0");
    if (capacity > 0 && c->ninst != 1) return fail(c, TLW_ENOTSUP, "the dt record needs a single-instance context");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    drop_graphs(c);
    (void)hipFree(c->d_dtlog);
    c->d_dtlog = nullptr;
    c->dtlog_cap = 0;
    if (capacity > 0) {
        HIPCHK(c, hipMalloc(&c->d_dtlog, sizeof(double) * (size_t)capacity));
        HIPCHK(c, hipMemset(c->d_dtlog, 0, sizeof(double) * (size_t)capacity));
        c->dtlog_cap = capacity;
    }
    return TLW_OK;
}

int tlw_get_dts(tlw_ctx* c, double* dts_host, int64_t n) {
    if (!c || (!dts_host && n > 0)) return fail(c, TLW_EINVAL, "null pointer");
    if (n < 0 || n > c->dtlog_cap) return fail(c, TLW_EINVAL, "n exceeds the dt record's capacity");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (n > 0) HIPCHK(c, hipMemcpy(dts_host, c->d_dtlog, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost));
    return TLW_OK;
}

int tlw_prescribe_dts(tlw_ctx* c, const double* dts_host, int64_t n) {
    if (!c || (!dts_host && n > 0)) return fail(c, TLW_EINVAL, "null pointer");
    if (n < 0) return fail(c, TLW_EINVAL, "n < 0");
    if (n > 0 && c->ninst != 1) return fail(c, TLW_ENOTSUP, "prescribed dts need a single-instance context");
    if (n > 0 && c->adaptive) return fail(c, TLW_ENOTSUP, "prescribed dts replace CFL steps, not adaptive ones");
    for (int64_t k = 0; k < n; ++k)
        if (!(dts_host[k] > 0.0) || !std::isfinite(dts_host[k]))
            return fail(c, TLW_EINVAL, "prescribed dts must be finite and > 0");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    drop_graphs(c);
    (void)hipFree(c->d_dtpresc);
    c->d_dtpresc = nullptr;
    c->dtpresc_n = 0;
    if (n > 0) {
        HIPCHK(c, hipMalloc(&c->d_dtpresc, sizeof(double) * (size_t)n));
        HIPCHK(c, hipMemcpy(c->d_dtpresc, dts_host, sizeof(double) * (size_t)n, hipMemcpyHostToDevice));
        c->dtpresc_n = n;
    }
    return TLW_OK;
}

int tlw_max_dt_noalloc(const double* u, const double* inv_jac, int32_t n_elements, int32_t polydeg, double gamma,
                       double* max_dt, void* stream) {
    if (!u || !inv_jac || !max_dt) {
        g_noctx_err = "tlw_max_dt_noalloc: null pointer";
        return TLW_EINVAL;
    }
    if (n_elements < 1 || (polydeg != 3 && polydeg != 4) || !(gamma > 1.0)) {
        g_noctx_err = "tlw_max_dt_noalloc: n_elements >= 1, polydeg 3 or 4 and gamma > 1 required";
        return TLW_EINVAL;
    }
    std::string err;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int rc = polydeg == 4 ? tlw_noalloc::launch<4>(u, inv_jac, n_elements, gamma, max_dt, s, err)
                                : tlw_noalloc::launch<3>(u, inv_jac, n_elements, gamma, max_dt, s, err);
    if (rc) g_noctx_err = "tlw_max_dt_noalloc: " + err;
    return rc;
}

int tlw_create_ensemble(tlw_ctx* a, tlw_ctx* b, tlw_ensemble** out) {
    if (!out) return TLW_EINVAL;
    *out = nullptr;
    if (!a || !b || a == b) {
        g_noctx_err = "tlw_create_ensemble: two distinct contexts required";
        return TLW_EINVAL;
    }
    tlw_ctx* c3 = a->N == 3 ? a : b;
    tlw_ctx* c4 = a->N == 3 ? b : a;
    if (c3->N != 3 || c4->N != 4) {
        g_noctx_err = "tlw_create_ensemble: one context of polydeg 3 and one of polydeg 4";
        return TLW_EINVAL;
    }
    for (tlw_ctx* c : {c3, c4})
        if (c->time_disc != TLW_LW || c->sc || c->parabolic || c->curved || c->has_boundary || c->positivity ||
            c->adaptive || c->d_dtpresc || c->d_dtlog || c->hconst.src_kind != TLW_SRC_NONE) {
            g_noctx_err = "tlw_create_ensemble: plain LW contexts on periodic TreeMeshes (no shock capturing, "
                          "parabolic terms, source terms, limiter, adaptive steps or dt record)";
            return TLW_ENOTSUP;
        }
    auto* e = new tlw_ensemble;
    e->c3 = c3;
    e->c4 = c4;
    e->stream = c3->stream;
    *out = e;
    return TLW_OK;
}

namespace {

int ens_fail(tlw_ensemble* e, int code, const std::string& m) {
    e->err = m;
    return code;
}

void ens_drop_graphs(tlw_ensemble* e) {
    for (auto& kv : e->graph_cache) (void)hipGraphExecDestroy(kv.second);
    e->graph_cache.clear();
}

// One LW step of both contexts: the volume, surface and finalize launches of
// each, merged into one launch apiece.  Tiles (16, 10) run 256-lane workgroups,
// (8, 5) 128-lane ones.
void enqueue_pair_step(tlw_ensemble* e) {
    tlw_ctx* c3 = e->c3;
    tlw_ctx* c4 = e->c4;
    const KArgs a3 = lw_step_args(c3), a4 = lw_step_args(c4);
    const dim3 gv(c3->vgrid + c4->vgrid);
    if (c3->ebv == 16)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_volume_pair<16, 10>), gv, dim3(VGeo<4, 10>::NT), 0, e->stream, a3, a4,
                           c3->vgrid);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_volume_pair<8, 5>), gv, dim3(VGeo<4, 5>::NT), 0, e->stream, a3, a4,
                           c3->vgrid);
    hipLaunchKernelGGL(k_surface_pair, dim3(c3->sgrid + c4->sgrid), dim3(kSurfBlock), 0, e->stream, a3, a4,
                       c3->sgrid);
    hipLaunchKernelGGL(k_finalize_pair, dim3(c3->ninst + c4->ninst), dim3(kFinalizeBlock), 0, e->stream,
                       fin_args(c3, RED_STEP, 0), fin_args(c4, RED_STEP, 0), c3->ninst);
}

}  // namespace

int tlw_ensemble_set_stream(tlw_ensemble* e, void* stream) {
    if (!e) return TLW_EINVAL;
    e->stream = static_cast<hipStream_t>(stream);
    e->c3->stream = e->stream;
    e->c4->stream = e->stream;
    return TLW_OK;
}

int tlw_ensemble_enable_graphs(tlw_ensemble* e, int32_t enable) {
    if (!e) return TLW_EINVAL;
    e->graphs = enable != 0;
    if (!e->graphs) ens_drop_graphs(e);
    return TLW_OK;
}

int tlw_ensemble_solve_init(tlw_ensemble* e, double t0, double t_final, double cfl3, double cfl4) {
    if (!e) return TLW_EINVAL;
    e->c3->stream = e->c4->stream = e->stream;
    int rc = tlw_solve_init(e->c3, t0, t_final, cfl3);
    if (rc) return ens_fail(e, rc, std::string("polydeg 3 context: ") + e->c3->err);
    rc = tlw_solve_init(e->c4, t0, t_final, cfl4);
    if (rc) return ens_fail(e, rc, std::string("polydeg 4 context: ") + e->c4->err);
    return TLW_OK;
}

int tlw_ensemble_advance(tlw_ensemble* e, int32_t max_steps, int32_t* finished) {
    if (!e) return TLW_EINVAL;
    if (max_steps < 0) return ens_fail(e, TLW_EINVAL, "max_steps < 0");
    tlw_ctx* c3 = e->c3;
    tlw_ctx* c4 = e->c4;
    const bool pair16 = c3->ebv == 16 && c4->ebv == 10, pair8 = c3->ebv == 8 && c4->ebv == 5;
    if (!pair16 && !pair8)
        return ens_fail(e, TLW_EINVAL, "tiles (elements per workgroup) must be (16, 10) or (8, 5)");
    c3->stream = c4->stream = e->stream;
    if (max_steps > 0) {
        if (e->graphs) {
            if (!e->stream) return ens_fail(e, TLW_EINVAL, "graph replay needs an explicit stream");
            if (e->graph_stream != e->stream) {
                ens_drop_graphs(e);
                e->graph_stream = e->stream;
            }
            // a tile change re-launches different kernels: key the cache on the pair too
            const int key = 2 * max_steps + (pair16 ? 1 : 0);
            auto it = e->graph_cache.find(key);
            if (it == e->graph_cache.end()) {
                hipGraph_t g = nullptr;
                hipError_t er = hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal);
                if (er == hipSuccess) {
                    for (int k = 0; k < max_steps; ++k) enqueue_pair_step(e);
                    er = hipStreamEndCapture(e->stream, &g);
                }
                hipGraphExec_t ge = nullptr;
                if (er == hipSuccess) er = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
                if (g) (void)hipGraphDestroy(g);
                if (er != hipSuccess) return ens_fail(e, TLW_EHIP, std::string("graph capture: ") + hipGetErrorString(er));
                it = e->graph_cache.emplace(key, ge).first;
            }
            hipError_t er = hipGraphLaunch(it->second, e->stream);
            if (er != hipSuccess) return ens_fail(e, TLW_EHIP, std::string("graph launch: ") + hipGetErrorString(er));
        } else {
            for (int k = 0; k < max_steps; ++k) enqueue_pair_step(e);
        }
    }
    hipError_t er = hipGetLastError();
    if (er != hipSuccess) return ens_fail(e, TLW_EHIP, std::string("kernel launch: ") + hipGetErrorString(er));
    bool fin = true, bad = false;
    for (tlw_ctx* c : {c3, c4}) {
        int rc = sync_state(c);
        if (rc) return ens_fail(e, rc, c->err);
        for (int k = 0; k < c->ninst; ++k) {
            c->host_cur[k] = c->h_state[k].cur;
            fin = fin && c->h_state[k].finished;
        }
        bad = bad || any_bad(c);
    }
    if (finished) *finished = fin ? 1 : 0;
    if (bad) return ens_fail(e, TLW_EADMISSIBLE, "non-admissible state (rho <= 0 or p <= 0)");
    return TLW_OK;
}

const char* tlw_ensemble_last_error(const tlw_ensemble* e) { return e ? e->err.c_str() : "null ensemble"; }

void tlw_ensemble_destroy(tlw_ensemble* e) {
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    ens_drop_graphs(e);
    delete e;
}

}  // extern "C"

namespace {

int create_common(int ninst, const tlw_mesh_desc* meshes, const tlw_basis_desc* basis, const tlw_solver_desc* solver,
                  tlw_ctx** out, const tlw_parabolic_desc* par, const tlw_curved_desc* cv, const tlw_mortar_desc* md,
                  const tlw_p4est_desc* pd) {
    if (!out) return TLW_EINVAL;
    *out = nullptr;
    if (!meshes || !basis || !solver) return TLW_EINVAL;
    if (basis->polydeg != 3 && basis->polydeg != 4) return TLW_ENOTSUP;
    if (!basis->nodes_host || !basis->weights_host || !basis->derivative_matrix_host || !basis->derivative_dhat_host)
        return TLW_EINVAL;
    if (solver->time_discretization != TLW_LW && solver->time_discretization != TLW_MDRK) return TLW_EINVAL;
    if (solver->surface_flux_order != TLW_FLUX_TREE && solver->surface_flux_order != TLW_FLUX_STRUCTURED)
        return TLW_EINVAL;
    if (solver->boundary_kind < TLW_BV_NONE || solver->boundary_kind > TLW_BV_SLIP_WALL_APPROXIMATE) return TLW_EINVAL;
    // the conditions on uⁿ with a normal direction: the curvilinear Euler mesh
    if ((solver->boundary_kind == TLW_BV_OUTFLOW || solver->boundary_kind == TLW_BV_SLIP_WALL_APPROXIMATE) &&
        (!cv || par))
        return TLW_ENOTSUP;
    // P4estMesh: the hyperbolic plain-FR kernels of the curvilinear mesh
    if (pd && (par || solver->shock_capturing || solver->boundary_kind == TLW_BV_SLIP_WALL || ninst != 1))
        return TLW_ENOTSUP;
    if (!(solver->gamma > 1.0)) return TLW_EINVAL;
    if (solver->source_terms < TLW_SRC_NONE || solver->source_terms > TLW_SRC_NAVIER_STOKES_CONVERGENCE_TEST)
        return TLW_EINVAL;
    // source terms: the Euler LW and MDRK kernels of the TreeMesh; the Navier-Stokes
    // convergence source: the Navier-Stokes MDRK kernels of the TreeMesh
    // (on a curvilinear StructuredMesh or a P4estMesh, whose volume kernels are the same: LW and MDRK
    // without shock capturing, structured/dg_2d.jl:1379-2105, structured/multi_derivative_2d.jl:82-445)
    if (solver->source_terms == TLW_SRC_CONVERGENCE_TEST && (solver->shock_capturing || par)) return TLW_ENOTSUP;
    if (solver->source_terms == TLW_SRC_NAVIER_STOKES_CONVERGENCE_TEST &&
        (!par || par->equations != TLW_PAR_NAVIER_STOKES || cv || pd || ninst != 1 ||
         (solver->time_discretization != TLW_MDRK && !par->walls)))
        return TLW_ENOTSUP;
    tlw_ctx* c = new tlw_ctx;
    c->N = basis->polydeg;
    c->n = c->N + 1;
    c->ebv = default_eb(c->N);
    c->p4est = pd != nullptr;
    c->time_disc = solver->time_discretization;
    c->flags = solver->flags;
    DevConst& h = c->hconst;
    const int n = c->n;
    for (int i = 0; i < n; ++i) {
        h.nodes[i] = basis->nodes_host[i];
        h.weights[i] = basis->weights_host[i];
        for (int k = 0; k < n; ++k) {
            h.D[i][k] = basis->derivative_matrix_host[i + n * k];
            h.Dhat[i][k] = basis->derivative_dhat_host[i + n * k];
        }
    }
    h.bnd_lo = 1.0 / h.weights[0];
    h.bnd_hi = 1.0 / h.weights[n - 1];
    mortar_operators(h, n);
    h.gamma = solver->gamma;
    for (int k = 0; k < 24; ++k) h.bc_par[k] = solver->boundary_params[k];
    h.bc_kind = solver->boundary_kind;
    h.flux_order = solver->surface_flux_order;
    h.src_kind = solver->source_terms;
    c->sc = solver->shock_capturing != 0;
    if (c->sc) {
        if (!solver->inverse_vandermonde_host || (solver->surface_flux_order != TLW_FLUX_TREE && !cv) ||
            !(solver->alpha_max >= 0.0) || !(solver->alpha_min >= 0.0)) {
            delete c;
            return TLW_EINVAL;
        }
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < n; ++k) h.Vinv[i][k] = solver->inverse_vandermonde_host[i + n * k];
        h.alpha_max = solver->alpha_max;
        h.alpha_min = solver->alpha_min;
        h.alpha_smooth = solver->alpha_smooth;
        if (solver->reconstruction != TLW_RECON_FIRST_ORDER && solver->reconstruction != TLW_RECON_MUSCL_HANCOCK) {
            delete c;
            return TLW_EINVAL;
        }
        h.recon = solver->reconstruction;
        if (h.recon == TLW_RECON_MUSCL_HANCOCK) {
            // compute_subcells (dg_2d.jl:1806-1824): faces xs[0] = -1, xs[i] = -1 + w_1 + ... + w_i
            // (summed left to right), extended nodes xe[0] = -1, xe[1..n] = nodes, xe[n+1] = 1
            double xs[9], xe[10];
            xs[0] = -1.0;
            for (int i = 1; i <= n; ++i) {
                double acc = h.weights[0];
                for (int k = 1; k < i; ++k) acc += h.weights[k];
                xs[i] = -1.0 + acc;
            }
            xe[0] = -1.0;
            for (int i = 1; i <= n; ++i) xe[i] = h.nodes[i - 1];
            xe[n + 1] = 1.0;
            for (int i = 0; i < n; ++i) {
                const int I = i + 1;  // 1-based subcell of node i
                const double d1 = xe[I] - xe[I - 1], d2 = xe[I + 1] - xe[I];
                h.mh_dx1[i] = d1;
                h.mh_dx2[i] = d2;
                // finite_differences (structured/dg_2d.jl:75-81); the end nodes (d1 or d2 = 0)
                // take zero slopes and never read these
                const bool inner = d1 > 0.0 && d2 > 0.0;
                h.mh_fa[i] = inner ? -(d2 / (d1 * (d1 + d2))) : 0.0;
                h.mh_fb[i] = inner ? (d2 - d1) / (d1 * d2) : 0.0;
                h.mh_fc[i] = inner ? (d1 / (d2 * (d1 + d2))) : 0.0;
                h.mh_xl[i] = xs[I - 1] - xe[I];
                h.mh_xr[i] = xs[I] - xe[I];
                h.mh_hw[i] = (i + 2 <= n) ? xs[i + 2] - xs[i + 1] : 1.0;
            }
        }
    }
    if (par) {
        // CompressibleNavierStokesDiffusion2D(equations; mu, Prandtl,
        // gradient_variables = GradientVariablesConservative()), constant μ
        const bool ok = par->equations == TLW_PAR_NAVIER_STOKES && par->gradient_variables == TLW_GRAD_CONSERVATIVE &&
                        std::isfinite(par->mu) && par->mu >= 0.0 && std::isfinite(par->prandtl) && par->prandtl > 0.0;
        // no-slip adiabatic or isothermal walls (LW or MDRK, TreeMesh, with the reflecting wall
        // of the hyperbolic part); wall temperatures must be finite and positive
        bool temps_ok = true;
        for (int k = 0; k < 4; ++k)
            if (par->walls && par->wall_isothermal[k])
                temps_ok = temps_ok && std::isfinite(par->wall_temperature[k]) && par->wall_temperature[k] > 0.0;
        const bool walls_ok = !par->walls || (!cv && solver->boundary_kind == TLW_BV_SLIP_WALL);
        if (!temps_ok) {
            delete c;
            return TLW_EINVAL;
        }
        const bool supported = !c->sc && ninst == 1 && walls_ok;
        if (!ok || !supported) {
            delete c;
            return ok ? TLW_ENOTSUP : TLW_EINVAL;
        }
        c->parabolic = true;
        h.visc_mu = par->mu;
        h.visc_Pr = par->prandtl;
        h.visc_kappa = solver->gamma * (1.0 / (solver->gamma - 1.0)) / par->prandtl;
        h.ns_walls = par->walls ? 1 : 0;
        // the wall kernel runs at one wave per SIMD: a tile of at most 4 waves
        if (h.ns_walls && c->ebv * c->n * c->n > 256) c->ebv = 10;  // polydeg 4 (polydeg 3 tiles fit)
        for (int k = 0; k < 8; ++k) h.wall_v[k] = par->walls ? par->wall_velocity[k] : 0.0;
        for (int k = 0; k < 4; ++k) {
            h.wall_iso[k] = (par->walls && par->wall_isothermal[k]) ? 1 : 0;
            h.wall_T[k] = h.wall_iso[k] ? par->wall_temperature[k] : 0.0;
        }

    }
    if (cv) {
        // StructuredMesh: LW or MDRK; shock capturing with the first-order subcell
        // scheme (MUSCL-Hancock is rejected by the reference on such meshes), or
        // compressible Navier-Stokes (periodic, checked at upload)
        if (ninst != 1 || (c->sc && h.recon != TLW_RECON_FIRST_ORDER)) {
            delete c;
            return TLW_ENOTSUP;
        }
        if (h.bc_kind == TLW_BV_SLIP_WALL) {  // the reference's curvilinear wall is another function
            delete c;
            return TLW_ENOTSUP;
        }
        c->curved = true;
        h.flux_order = TLW_FLUX_STRUCTURED;
    }
    auto bail = [&](int rc) {
        tlw_destroy(c);
        return rc;
    };
    if (hipMalloc(&c->d_const, sizeof(DevConst)) != hipSuccess) return bail(TLW_EHIP);
    if (hipMemcpy(c->d_const, &h, sizeof(DevConst), hipMemcpyHostToDevice) != hipSuccess) return bail(TLW_EHIP);
    int rc = upload_meshes(c, meshes, ninst, cv, md, pd);
    if (rc) return bail(rc);
    // Trixi's alpha smoothing on a StructuredMesh assumes a periodic mesh
    if (c->curved && c->sc && h.alpha_smooth && c->has_boundary) return bail(TLW_ENOTSUP);
    *out = c;
    return TLW_OK;
}

}  // namespace
