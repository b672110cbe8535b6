// C entry points of libjsso_hip.so and the host-side symbolic plan.
//
// Block pattern.  The reference emits, for every element, the full dense
// meshgrid of its DOFs (element.py:141-149, 1097-1106) plus a (0,0) zero seed
// (assemblemodel.py:202-204) and sums duplicates (assemblemodel.py:160-162), so
// its scalar pattern is exactly the 6x6-block pattern of the node adjacency
// graph (nodes sharing an element), which is what is built here, one sorted
// block row per node.  Every node also gets its diagonal block (the solver
// needs it; for a node outside all elements this adds zeros the reference
// would not store, and such a node makes K singular in both).
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <numeric>

#include "jsso_internal.hpp"

namespace jsso {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

template <typename T>
static int upload(Plan* p, T** dst, const std::vector<T>& src) {
    int rc = dalloc(p, dst, src.size());
    if (rc) return rc;
    if (!src.empty()) JSSO_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return JSSO_OK;
}

// Symmetric sliced ELL of the off-diagonal blocks of a block-CSR pattern
// (sorted columns, diagonal present in every row) over packed structures
// (structure s owns rows [node_off[s], node_off[s+1]); an empty range gives no
// slice).  A slice is <= ROWS_PER_WG block rows of one structure; lane l is
// row row0 + l.  Only the upper blocks (i < j) are stored: value slot u of a
// slice holds the u-th upper block of every row.  The column list of a slice
// has nup upper slots (column j, block stored in this slice) followed by nlo
// lower slots (column j < i; the block is the transpose of the upper block
// (j, i), found through its value position vpos = value slot * 64 + lane).
// Padding: upper slots hold zero values with the row's own column; lower slots
// point at an all-zero value slot appended after the last one.
int build_ell(int32_t n_sys, const int32_t* node_off, const std::vector<int32_t>& rowptr,
              const std::vector<int32_t>& colind, EllPattern& out) {
    const int32_t n = node_off[n_sys];
    const int64_t nnzb = colind.size();
    std::vector<int32_t> diag(n);
    for (int i = 0; i < n; ++i) {
        auto b = colind.begin() + rowptr[i], e = colind.begin() + rowptr[i + 1];
        auto it = std::lower_bound(b, e, i);
        JSSO_REQUIRE(it != e && *it == i, JSSO_EINVAL, "block row %d has no diagonal block", i);
        diag[i] = (int32_t)(it - colind.begin());
    }
    // position in the block CSR of the transposed block (j, i) of every block (i, j)
    std::vector<int32_t> tr(nnzb);
    for (int i = 0; i < n; ++i)
        for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
            const int j = colind[k];
            const int32_t* b = colind.data() + rowptr[j];
            const int32_t* e = colind.data() + rowptr[j + 1];
            tr[k] = (int32_t)(std::lower_bound(b, e, i) - colind.data());
        }
    std::vector<int32_t>& wg_sys = out.wg_sys;
    std::vector<int32_t>& wg_row0 = out.wg_row0;
    std::vector<int32_t>& wg_row1 = out.wg_row1;
    std::vector<int32_t>& sys_wg_off = out.sys_wg_off;
    wg_sys.clear();
    wg_row0.clear();
    wg_row1.clear();
    sys_wg_off.assign(n_sys + 1, 0);
    for (int s = 0; s < n_sys; ++s) {
        for (int r0 = node_off[s]; r0 < node_off[s + 1]; r0 += ROWS_PER_WG) {
            wg_sys.push_back(s);
            wg_row0.push_back(r0);
            wg_row1.push_back(std::min(r0 + ROWS_PER_WG, node_off[s + 1]));
        }
        sys_wg_off[s + 1] = (int32_t)wg_sys.size();
    }
    const int n_wg = (int)wg_sys.size();
    out.n_wg = n_wg;
    std::vector<int32_t>& ell_col = out.ell_col;   // per column position
    std::vector<int32_t>& ell_vpos = out.ell_vpos;
    std::vector<int32_t>& ell_src = out.ell_src;   // per value position
    ell_col.clear();
    ell_vpos.clear();
    ell_src.clear();
    std::vector<int32_t> pos_off(n_wg), val_off(n_wg), nupv(n_wg), nlov(n_wg);
    std::vector<int64_t> vpos_of(nnzb, -1);
    int64_t npos = 0, nval = 0;
    for (int w = 0; w < n_wg; ++w) {
        int nup = 0, nlo = 0;
        for (int i = wg_row0[w]; i < wg_row1[w]; ++i) {
            nup = std::max(nup, rowptr[i + 1] - diag[i] - 1);
            nlo = std::max(nlo, diag[i] - rowptr[i]);
        }
        JSSO_REQUIRE(nup < 65536 && nlo < 65536, JSSO_EINVAL, "a block row has more than 65535 neighbours");
        pos_off[w] = (int32_t)npos;
        val_off[w] = (int32_t)nval;
        nupv[w] = nup;
        nlov[w] = nlo;
        ell_col.resize((size_t)(npos + (int64_t)nup * ROWS_PER_WG));
        ell_vpos.resize(ell_col.size(), 0);
        ell_src.resize((size_t)(nval + (int64_t)nup * ROWS_PER_WG), -1);
        for (int u = 0; u < nup; ++u)
            for (int l = 0; l < ROWS_PER_WG; ++l) {
                const int i = wg_row0[w] + l;
                const int64_t pos = npos + (int64_t)u * ROWS_PER_WG + l;
                const int64_t vp = nval + (int64_t)u * ROWS_PER_WG + l;
                int32_t c = i < wg_row1[w] ? i : wg_row0[w];
                if (i < wg_row1[w] && u < rowptr[i + 1] - diag[i] - 1) {
                    const int k = diag[i] + 1 + u;
                    c = colind[k];
                    ell_src[vp] = k;
                    vpos_of[k] = vp;
                }
                ell_col[pos] = c;
                ell_vpos[pos] = (int32_t)vp;
            }
        npos += (int64_t)nup * ROWS_PER_WG;
        nval += (int64_t)nup * ROWS_PER_WG;
        JSSO_REQUIRE(npos < INT32_MAX / 2 && nval < INT32_MAX / 36, JSSO_EINVAL,
                     "plan too large for 32-bit sliced-ELL offsets");
    }
    // lower slots, once every upper block has its value position (the zero slot
    // comes after the last value slot)
    const int64_t zero_vp = nval;
    std::vector<int32_t> lo_off(n_wg);
    {
        int64_t total = npos;
        for (int w = 0; w < n_wg; ++w) total += (int64_t)nlov[w] * ROWS_PER_WG;
        JSSO_REQUIRE(total < INT32_MAX, JSSO_EINVAL, "plan too large for 32-bit sliced-ELL offsets");
        ell_col.resize((size_t)total);
        ell_vpos.resize((size_t)total, 0);
    }
    for (int w = 0; w < n_wg; ++w) {
        lo_off[w] = (int32_t)npos;
        const int nlo = nlov[w];
        for (int u = 0; u < nlo; ++u)
            for (int l = 0; l < ROWS_PER_WG; ++l) {
                const int i = wg_row0[w] + l;
                const int64_t pos = npos + (int64_t)u * ROWS_PER_WG + l;
                int32_t c = i < wg_row1[w] ? i : wg_row0[w];
                int64_t vp = zero_vp;
                if (i < wg_row1[w] && u < diag[i] - rowptr[i]) {
                    const int k = rowptr[i] + u;
                    c = colind[k];
                    vp = vpos_of[tr[k]];
                }
                ell_col[pos] = c;
                ell_vpos[pos] = (int32_t)vp;
            }
        npos += (int64_t)nlo * ROWS_PER_WG;
    }
    out.ell_slots = nval + ROWS_PER_WG;   // value positions incl. the zero slot
    // per block of the CSR: 2 * value position of its stored block, + 1 when
    // the block is the transpose of the stored one; -1 for diagonal blocks
    out.kref.assign(nnzb, -1);
    for (int i = 0; i < n; ++i)
        for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
            if (k == diag[i]) continue;
            out.kref[k] = k > diag[i] ? (int32_t)(2 * vpos_of[k]) : (int32_t)(2 * vpos_of[tr[k]] + 1);
        }
    // per slice, three int4: {sys, row0, row1, first value position},
    // {first upper column position, upper slots, first lower column position, lower slots},
    // {first slice of sys, end slice of sys, 0, 0}
    std::vector<int32_t>& meta = out.meta;
    meta.assign(12 * (size_t)n_wg, 0);
    for (int w = 0; w < n_wg; ++w) {
        int32_t* m = &meta[12 * (size_t)w];
        m[0] = wg_sys[w];
        m[1] = wg_row0[w];
        m[2] = wg_row1[w];
        m[3] = val_off[w];
        m[4] = pos_off[w];
        m[5] = nupv[w];
        m[6] = lo_off[w];
        m[7] = nlov[w];
        m[8] = sys_wg_off[wg_sys[w]];
        m[9] = sys_wg_off[wg_sys[w] + 1];
    }
    return JSSO_OK;
}

static int build_plan(Plan* p, int32_t n_sys, const int32_t* node_off, int32_t n_beam, const int32_t* cb,
                      int32_t n_quad, const int32_t* cq, int32_t n_cons, const int32_t* cons) {
    JSSO_REQUIRE(n_sys >= 1 && node_off, JSSO_EINVAL, "n_sys must be >= 1 with node offsets");
    JSSO_REQUIRE(n_beam >= 0 && n_quad >= 0 && n_cons >= 0, JSSO_EINVAL, "negative element/constraint count");
    JSSO_REQUIRE(node_off[0] == 0, JSSO_EINVAL, "node_off[0] must be 0");
    for (int s = 0; s < n_sys; ++s)
        JSSO_REQUIRE(node_off[s + 1] > node_off[s], JSSO_EINVAL, "structure %d has no nodes", s);
    const int32_t n = node_off[n_sys];
    p->n_sys = n_sys;
    p->n_nodes = n;
    p->n_beam = n_beam;
    p->n_quad = n_quad;
    p->n_cons = n_cons;
    p->h_node_off.assign(node_off, node_off + n_sys + 1);
    std::vector<int32_t> sys_of(n);
    for (int s = 0; s < n_sys; ++s)
        for (int i = node_off[s]; i < node_off[s + 1]; ++i) sys_of[i] = s;
    auto check_el = [&](const int32_t* c, int nen, int e, const char* kind) -> int {
        for (int a = 0; a < nen; ++a)
            JSSO_REQUIRE(c[a] >= 0 && c[a] < n, JSSO_EINVAL, "%s %d: node %d out of range [0,%d)", kind, e, c[a], n);
        for (int a = 1; a < nen; ++a)
            JSSO_REQUIRE(sys_of[c[a]] == sys_of[c[0]], JSSO_EINVAL, "%s %d spans two structures", kind, e);
        return JSSO_OK;
    };
    for (int e = 0; e < n_beam; ++e) {
        int rc = check_el(cb + 2 * e, 2, e, "beam");
        if (rc) return rc;
    }
    for (int e = 0; e < n_quad; ++e) {
        int rc = check_el(cq + 4 * e, 4, e, "quad");
        if (rc) return rc;
    }
    // constrained DOFs
    std::vector<uint8_t> fixed(6 * (size_t)n, 0);
    for (int k = 0; k < n_cons; ++k) {
        const int64_t d = cons[k];
        JSSO_REQUIRE(d >= 0 && d < 6 * (int64_t)n, JSSO_EINVAL, "constrained DOF %lld out of range", (long long)d);
        JSSO_REQUIRE(!fixed[d], JSSO_EINVAL,
                     "constrained DOF %lld listed twice (the reference's K_aug would be singular)", (long long)d);
        fixed[d] = 1;
    }
    // node -> element incidence (element order: beams, then quads)
    std::vector<int32_t> inc_ptr(n + 1, 0);
    for (int e = 0; e < n_beam; ++e)
        for (int a = 0; a < 2; ++a) inc_ptr[cb[2 * e + a] + 1]++;
    for (int e = 0; e < n_quad; ++e)
        for (int a = 0; a < 4; ++a) inc_ptr[cq[4 * e + a] + 1]++;
    for (int i = 0; i < n; ++i) inc_ptr[i + 1] += inc_ptr[i];
    std::vector<int32_t> inc(inc_ptr[n]), inc_pos(inc_ptr.begin(), inc_ptr.end() - 1);
    std::vector<int8_t> inc_a(inc_ptr[n]);
    for (int e = 0; e < n_beam; ++e)
        for (int a = 0; a < 2; ++a) {
            const int nd = cb[2 * e + a];
            inc_a[inc_pos[nd]] = (int8_t)a;
            inc[inc_pos[nd]++] = e;
        }
    for (int e = 0; e < n_quad; ++e)
        for (int a = 0; a < 4; ++a) {
            const int nd = cq[4 * e + a];
            inc_a[inc_pos[nd]] = (int8_t)a;
            inc[inc_pos[nd]++] = n_beam + e;
        }
    // block rows
    std::vector<int32_t>& rowptr = p->h_rowptr;
    std::vector<int32_t>& colind = p->h_colind;
    rowptr.assign(n + 1, 0);
    colind.clear();
    std::vector<int32_t> row;
    for (int i = 0; i < n; ++i) {
        row.clear();
        row.push_back(i);
        for (int k = inc_ptr[i]; k < inc_ptr[i + 1]; ++k) {
            const int el = inc[k];
            if (el < n_beam) {
                row.push_back(cb[2 * el]);
                row.push_back(cb[2 * el + 1]);
            } else {
                const int e = el - n_beam;
                for (int a = 0; a < 4; ++a) row.push_back(cq[4 * e + a]);
            }
        }
        std::sort(row.begin(), row.end());
        row.erase(std::unique(row.begin(), row.end()), row.end());
        colind.insert(colind.end(), row.begin(), row.end());
        rowptr[i + 1] = (int32_t)colind.size();
    }
    const int64_t nnzb = colind.size();
    p->nnzb = nnzb;
    auto find_block = [&](int i, int j) -> int64_t {
        auto b = colind.begin() + rowptr[i], e = colind.begin() + rowptr[i + 1];
        auto it = std::lower_bound(b, e, j);
        return (it != e && *it == j) ? (int64_t)(it - colind.begin()) : -1;
    };
    std::vector<int32_t> browof(nnzb), diag(n);
    for (int i = 0; i < n; ++i) {
        for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) browof[k] = i;
        diag[i] = (int32_t)find_block(i, i);
    }
    // block -> element-block contributions, element order
    p->n_contrib = 4 * (int64_t)n_beam + 16 * (int64_t)n_quad;
    std::vector<int32_t> cptr(nnzb + 1, 0), cblk(p->n_contrib);
    for (int e = 0; e < n_beam; ++e)
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) {
                const int64_t k = find_block(cb[2 * e + a], cb[2 * e + b]);
                cblk[(int64_t)e * 4 + a * 2 + b] = (int32_t)k;
                cptr[k + 1]++;
            }
    for (int e = 0; e < n_quad; ++e)
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) {
                const int64_t k = find_block(cq[4 * e + a], cq[4 * e + b]);
                cblk[4 * (int64_t)n_beam + (int64_t)e * 16 + a * 4 + b] = (int32_t)k;
                cptr[k + 1]++;
            }
    for (int64_t k = 0; k < nnzb; ++k) cptr[k + 1] += cptr[k];
    std::vector<int32_t> csrc(p->n_contrib), cfill(cptr.begin(), cptr.end() - 1);
    for (int64_t src = 0; src < p->n_contrib; ++src) csrc[cfill[cblk[src]]++] = (int32_t)src;
    // node -> adjoint gradient buffer offsets
    std::vector<int32_t> nptr(n + 1, 0), nsrc(inc_ptr[n]);
    for (int i = 0; i < n; ++i) {
        nptr[i + 1] = inc_ptr[i + 1];
        for (int k = inc_ptr[i]; k < inc_ptr[i + 1]; ++k) {
            const int el = inc[k];
            const int a = inc_a[k];
            nsrc[k] = el < n_beam ? el * BDIR + 3 * a : n_beam * BDIR + (el - n_beam) * QDIR + 3 * a;
        }
    }
    // off-diagonal solver pattern
    p->nnzb_off = nnzb - n;
    EllPattern ell;
    int rc = build_ell(n_sys, node_off, rowptr, colind, ell);
    if (rc) return rc;
    p->n_wg = ell.n_wg;
    p->h_sys_wg_off = ell.sys_wg_off;
    // (structures below 64 slices are latency-bound per launch wherever they run;
    // JSSO_SLICE_SPREAD=0 / 1 forces the contiguous map / the runs for any batch:
    // the results are the same bits either way, the tests compare them)
    int max_wg = 0;
    for (int s = 0; s < n_sys; ++s) max_wg = std::max(max_wg, ell.sys_wg_off[s + 1] - ell.sys_wg_off[s]);
    p->slice_spread = n_sys > 1 && max_wg >= 64;
    if (const char* e = getenv("JSSO_SLICE_SPREAD")) p->slice_spread = e[0] == '1' && n_sys > 1;
    p->ell_slots = ell.ell_slots;
    p->h_kref.swap(ell.kref);
    std::vector<int32_t>& meta = ell.meta;
    std::vector<int32_t>& ell_col = ell.ell_col;
    std::vector<int32_t>& ell_vpos = ell.ell_vpos;
    std::vector<int32_t>& ell_src = ell.ell_src;
    std::vector<int32_t>& wg_sys = ell.wg_sys;
    std::vector<int32_t>& wg_row0 = ell.wg_row0;
    std::vector<int32_t>& wg_row1 = ell.wg_row1;
    p->h_slice_meta.swap(meta);
    p->h_ell_col.swap(ell_col);
    p->h_ell_vpos.swap(ell_vpos);
    p->h_ell_src.swap(ell_src);
    // position of the transposed block of every block (the pattern is structurally symmetric)
    p->h_tpos.resize(nnzb);
    for (int64_t k = 0; k < nnzb; ++k) p->h_tpos[k] = (int32_t)find_block(colind[k], browof[k]);
    p->h_browof.swap(browof);
    p->h_diag.swap(diag);
    p->h_cptr.swap(cptr);
    p->h_csrc.swap(csrc);
    p->h_nptr.swap(nptr);
    p->h_nsrc.swap(nsrc);
    p->h_cnct_beam.assign(cb, cb + 2 * (size_t)n_beam);
    p->h_cnct_quad.assign(cq, cq + 4 * (size_t)n_quad);
    p->h_wg_sys.swap(wg_sys);
    p->h_wg_row0.swap(wg_row0);
    p->h_wg_row1.swap(wg_row1);
    p->h_fixed.swap(fixed);
    return JSSO_OK;
}

// Upload the symbolic arrays and allocate the numeric buffers (first use).
static int ensure_device(Plan* p) {
    if (p->on_device) return JSSO_OK;
    int rc = 0;
    if ((rc = upload(p, &p->d_rowptr, p->h_rowptr)) || (rc = upload(p, &p->d_colind, p->h_colind)) ||
        (rc = upload(p, &p->d_browof, p->h_browof)) || (rc = upload(p, &p->d_diag, p->h_diag)) ||
        (rc = upload(p, &p->d_cptr, p->h_cptr)) ||
        (rc = upload(p, &p->d_csrc, p->h_csrc)) || (rc = upload(p, &p->d_nptr, p->h_nptr)) ||
        (rc = upload(p, &p->d_nsrc, p->h_nsrc)) || (rc = upload(p, &p->d_cnct_beam, p->h_cnct_beam)) ||
        (rc = upload(p, &p->d_cnct_quad, p->h_cnct_quad)) || (rc = upload(p, &p->d_fixed, p->h_fixed)) ||
        (rc = upload(p, &p->d_wg_sys, p->h_wg_sys)) || (rc = upload(p, &p->d_wg_row0, p->h_wg_row0)) ||
        (rc = upload(p, &p->d_wg_row1, p->h_wg_row1)) || (rc = upload(p, &p->d_sys_wg_off, p->h_sys_wg_off)) ||
        (rc = upload(p, &p->d_ell_vpos, p->h_ell_vpos)) || (rc = upload(p, &p->d_ell_col, p->h_ell_col)) ||
        (rc = upload(p, &p->d_slice_meta, p->h_slice_meta)) ||
        (rc = upload(p, &p->d_ell_src, p->h_ell_src)) || (rc = upload(p, &p->d_tpos, p->h_tpos)))
        return rc;
    const size_t n = p->n_nodes;
    if ((rc = dalloc(p, &p->d_ke, (size_t)p->n_contrib * 36)) || (rc = dalloc(p, &p->d_K, (size_t)p->nnzb * 36)) ||
        (rc = dalloc(p, &p->d_Kc, (size_t)p->nnzb * 36)) || (rc = dalloc(p, &p->d_Linv, n * 36)) ||
        (rc = dalloc(p, &p->d_Minv, n * 36)) || (rc = dalloc(p, &p->d_status, 1)) ||
        (rc = dalloc(p, &p->d_Ke64, (size_t)p->ell_slots * 36)) || (rc = dalloc(p, &p->d_Ke32, (size_t)p->ell_slots * 36)) ||
        (rc = dalloc(p, &p->d_Ke16, (size_t)p->ell_slots * 36)) ||
        (rc = dalloc(p, &p->d_crds, n * 3)))
        return rc;
    JSSO_HIP(hipMemset(p->d_status, 0, sizeof(int32_t)));
    // padding slots of the ELL arrays stay zero
    if (p->ell_slots) {
        JSSO_HIP(hipMemset(p->d_Ke64, 0, (size_t)p->ell_slots * 36 * sizeof(double)));
        JSSO_HIP(hipMemset(p->d_Ke32, 0, (size_t)p->ell_slots * 36 * sizeof(float)));
        JSSO_HIP(hipMemset(p->d_Ke16, 0, (size_t)p->ell_slots * 36 * sizeof(uint16_t)));
    }
    p->on_device = true;
    return JSSO_OK;
}

}  // namespace jsso

using namespace jsso;

struct jsso_plan {
    Plan impl;
};

static Plan* P(jsso_plan* p) { return &p->impl; }
static const Plan* P(const jsso_plan* p) { return &p->impl; }

extern "C" {

int jsso_version(void) { return 1; }

int jsso_last_error(char* buf, size_t n) {
    if (!buf || n == 0) return JSSO_EINVAL;
    snprintf(buf, n, "%s", g_err);
    return JSSO_OK;
}

int jsso_plan_create(int32_t n_sys, const int32_t* h_node_off, int32_t n_beam, const int32_t* h_cnct_beam,
                     int32_t n_quad, const int32_t* h_cnct_quad, int32_t n_cons, const int32_t* h_cons_dofs,
                     jsso_plan** out) {
    if (!out) {
        set_error("out is NULL");
        return JSSO_EINVAL;
    }
    *out = nullptr;
    if ((n_beam > 0 && !h_cnct_beam) || (n_quad > 0 && !h_cnct_quad) || (n_cons > 0 && !h_cons_dofs)) {
        set_error("NULL connectivity or constraint array with a positive count");
        return JSSO_EINVAL;
    }
    jsso_plan* p = new jsso_plan();
    const int rc = build_plan(P(p), n_sys, h_node_off, n_beam, h_cnct_beam, n_quad, h_cnct_quad, n_cons,
                              h_cons_dofs);
    if (rc) {
        Plan* q = P(p);
        for (void* a : q->allocs) (void)hipFree(a);
        delete p;
        return rc;
    }
    *out = p;
    return JSSO_OK;
}

int jsso_plan_destroy(jsso_plan* plan) {
    if (!plan) return JSSO_OK;
    Plan* q = P(plan);
    amg_destroy(q);
    for (auto& row : q->cg_graph)
        for (auto& pair : row)
            for (auto& g : pair)
                if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (q->capture_stream) (void)hipStreamDestroy(q->capture_stream);
    for (auto& e : q->flag_ev)
        if (e) (void)hipEventDestroy(e);
    if (q->h_flags) (void)hipHostFree(q->h_flags);
    if (q->h_rel) (void)hipHostFree(q->h_rel);
    for (void* a : q->allocs) (void)hipFree(a);
    for (auto& b : q->ws)
        if (b.p) (void)hipFree(b.p);
    if (q->d_dense) (void)hipFree(q->d_dense);
    delete plan;
    return JSSO_OK;
}

int jsso_plan_sizes(const jsso_plan* plan, int64_t* h_sizes) {
    if (!plan || !h_sizes) {
        set_error("NULL argument");
        return JSSO_EINVAL;
    }
    const Plan* p = P(plan);
    h_sizes[JSSO_SZ_NODES] = p->n_nodes;
    h_sizes[JSSO_SZ_NNZB] = p->nnzb;
    h_sizes[JSSO_SZ_BEAMS] = p->n_beam;
    h_sizes[JSSO_SZ_QUADS] = p->n_quad;
    h_sizes[JSSO_SZ_SYSTEMS] = p->n_sys;
    h_sizes[JSSO_SZ_CONTRIB] = p->n_contrib;
    return JSSO_OK;
}

int jsso_plan_pattern(const jsso_plan* plan, int32_t* h_rowptr, int32_t* h_colind) {
    if (!plan || !h_rowptr || !h_colind) {
        set_error("NULL argument");
        return JSSO_EINVAL;
    }
    const Plan* p = P(plan);
    std::copy(p->h_rowptr.begin(), p->h_rowptr.end(), h_rowptr);
    std::copy(p->h_colind.begin(), p->h_colind.end(), h_colind);
    return JSSO_OK;
}

int jsso_element_ke_beam(int32_t n, const double* d_crds, const double* d_props, double* d_ke, int32_t* d_flags,
                         void* stream) {
    if (n < 0 || (n > 0 && (!d_crds || !d_props || !d_ke))) {
        set_error("invalid arguments");
        return JSSO_EINVAL;
    }
    return launch_ke_beam(n, d_crds, nullptr, nullptr, d_props, d_ke, d_flags, nullptr, (hipStream_t)stream);
}

int jsso_element_ke_quad(int32_t n, const double* d_crds, const double* d_props, double* d_ke, void* stream) {
    if (n < 0 || (n > 0 && (!d_crds || !d_props || !d_ke))) {
        set_error("invalid arguments");
        return JSSO_EINVAL;
    }
    return launch_ke_quad(n, d_crds, nullptr, nullptr, d_props, d_ke, (hipStream_t)stream);
}

int jsso_element_beam_local(int32_t n, const double* d_crds, const double* d_props, double* d_T, double* d_klocal,
                            void* stream) {
    if (n < 0 || (n > 0 && (!d_crds || (d_klocal && !d_props)))) {
        set_error("invalid arguments");
        return JSSO_EINVAL;
    }
    return launch_beam_local(n, d_crds, d_props, d_T, d_klocal, (hipStream_t)stream);
}

int jsso_element_quad_local(int32_t n, const double* d_crds, const double* d_props, double* d_xy, double* d_normals,
                            double* d_T, double* d_klocal, void* stream) {
    if (n < 0 || (n > 0 && (!d_crds || (d_klocal && !d_props)))) {
        set_error("invalid arguments");
        return JSSO_EINVAL;
    }
    return launch_quad_local(n, d_crds, d_props, d_xy, d_normals, d_T, d_klocal, (hipStream_t)stream);
}

int jsso_element_quad_gauss(int32_t n, const double* d_xy, const double* d_props, double r, double s, double* d_J,
                            double* d_Bk, double* d_Bg, double* d_Bm, double* d_Cb, double* d_Cs, double* d_Cm,
                            void* stream) {
    const bool geo = d_J || d_Bk || d_Bg || d_Bm, mat = d_Cb || d_Cs || d_Cm;
    if (n < 0 || (n > 0 && ((geo && !d_xy) || (mat && !d_props)))) {
        set_error("invalid arguments");
        return JSSO_EINVAL;
    }
    return launch_quad_gauss(n, geo ? d_xy : nullptr, mat ? d_props : nullptr, r, s, d_J, d_Bk, d_Bg, d_Bm, d_Cb,
                             d_Cs, d_Cm, (hipStream_t)stream);
}

int jsso_assemble(jsso_plan* plan, const double* d_crds, const double* d_prop_beam, const double* d_prop_quad,
                  double* d_K_out, void* stream) {
    if (!plan || !d_crds) {
        set_error("NULL plan or coordinates");
        return JSSO_EINVAL;
    }
    Plan* p = P(plan);
    if ((p->n_beam > 0 && !d_prop_beam) || (p->n_quad > 0 && !d_prop_quad)) {
        set_error("missing element properties");
        return JSSO_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    int rc0 = ensure_device(p);
    if (rc0) return rc0;
    JSSO_HIP(hipMemsetAsync(p->d_status, 0, sizeof(int32_t), s));
    p->status_cache = -1;
    std::pair<hipEvent_t, hipEvent_t>* ev = nullptr;
    timing_begin_slot(s, ev, JSSO_TIMING_ELEMENT);
    int rc = launch_ke_beam(p->n_beam, nullptr, p->d_cnct_beam, d_crds, d_prop_beam, p->d_ke, nullptr, p->d_status, s);
    if (rc) return rc;
    double* ke_q = p->d_ke + (size_t)p->n_beam * 4 * 36;
    if (p->n_quad > 0 && p->ke_epw <= 0 && (rc = tune_ke_quad(p, d_crds, d_prop_quad, ke_q, s, p->ke_epw))) return rc;
    rc = launch_ke_quad(p->n_quad, nullptr, p->d_cnct_quad, d_crds, d_prop_quad, ke_q, s, p->ke_epw, p->d_status);
    if (rc) return rc;
    timing_end(s, ev);
    timing_begin_slot(s, ev, JSSO_TIMING_GATHER);
    rc = launch_assemble(p, s);
    if (rc) return rc;
    timing_end(s, ev);
    JSSO_HIP(hipMemcpyAsync(p->d_crds, d_crds, (size_t)p->n_nodes * 3 * sizeof(double), hipMemcpyDeviceToDevice, s));
    p->amg_valid = false;
    if (d_K_out && p->nnzb) {
        if ((rc = launch_full_blocks(p, s))) return rc;
        JSSO_HIP(hipMemcpyAsync(d_K_out, p->d_K, (size_t)p->nnzb * 36 * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    p->assembled = true;
    return JSSO_OK;
}

// K u = f (or K^T u = f) for a plan whose K is not symmetric.  The reference
// factorises K_aug (and K_aug^T for the adjoint) with SuperLU
// (solver.py:195-209,236).  Here: the dense LU (jsso_dense_solve) up to
// NONSYM_DENSE_MAX_DOF DOF or with members parallel to global Y; otherwise a
// defect correction preconditioned by the batched solver on the symmetric part
// S = (K + K^T)/2 (which the assembly writes into the solver copies of such a
// plan):  r = f - op(K) u;  S du = r (inner AMG-CG, relative tolerance 1e-3);
// u += du, until max ||r|| / ||f|| <= tol over the systems.  It converges when
// the spectral radius of I - S^-1 K is below 1: 0.04-0.14 for MITC-4 shells with
// kx_mod/ky_mod from 1.2/0.8 to 2/0.5; the transformation of Y-parallel members
// leaves S indefinite, so those plans stay on the dense LU.
static int solve_nonsymmetric(Plan* p, const double* d_f, double* d_u, int nrhs, int precision, double tol,
                              int maxit, jsso_info* h_info, bool transpose, hipStream_t s, int32_t status) {
    int rc = launch_full_blocks(p, s);
    if (rc) return rc;
    const int64_t ndof = 6 * (int64_t)p->n_nodes;
    const bool dense = p->nonsym_route == 1 ||
                       (p->nonsym_route == 0 && ((status & STATUS_Y_BEAM) || ndof <= NONSYM_DENSE_MAX_DOF));
    if (dense) {
        if ((rc = dense_solve(p, p->d_K, d_f, d_u, nrhs, true, transpose, s))) return rc;
        if (h_info) {
            *h_info = jsso_info{};
            h_info->converged = p->n_sys * nrhs;
            h_info->n_systems = p->n_sys * nrhs;
        }
        return JSSO_OK;
    }
    JSSO_REQUIRE(!(status & STATUS_Y_BEAM), JSSO_EUNSUPPORTED,
                 "members parallel to global Y (element.py:92-94) leave the symmetric part of K indefinite: the "
                 "defect correction cannot converge; use the dense route (jsso_plan_nonsymmetric_route 0 or 1)");
    const int nsys = p->n_sys * nrhs;
    const size_t vb = (size_t)ndof * nrhs * sizeof(double);
    const size_t need = 2 * vb + 2 * (size_t)nsys * sizeof(double);
    if (p->ws.size() <= WS_DEFECT_SLOT) p->ws.resize(WS_DEFECT_SLOT + 1);
    DBuf& g = p->ws[WS_DEFECT_SLOT];
    if (g.bytes < need) {
        if (g.p) JSSO_HIP(hipFree(g.p));
        g.p = nullptr;
        g.bytes = 0;
        hipError_t e = hipMalloc(&g.p, need);
        JSSO_REQUIRE(e == hipSuccess, JSSO_ENOMEM, "defect-correction vectors hipMalloc(%zu): %s", need,
                     hipGetErrorString(e));
        g.bytes = need;
    }
    double* r = (double*)g.p;
    double* du = r + (size_t)ndof * nrhs;
    double* d_rr = du + (size_t)ndof * nrhs;
    std::vector<double> rr(nsys), ff(nsys);
    const int slot = precision == JSSO_PREC_MIXED ? PS_F32 : precision == JSSO_PREC_MIXED_V64 ? PS_F32_V64 : PS_F64;
    JSSO_HIP(hipMemsetAsync(d_u, 0, vb, s));
    int inner_its = 0, outer = 0, conv = 0;
    double relmax = 0.0, prev = 0.0;
    for (;; ++outer) {
        if ((rc = launch_defect_residual(p, p->d_K, d_f, d_u, nrhs, transpose, r, s))) return rc;
        if ((rc = launch_compliance(p, r, r, nrhs, 1.0, d_rr, s))) return rc;
        JSSO_HIP(hipMemcpyAsync(rr.data(), d_rr, nsys * sizeof(double), hipMemcpyDeviceToHost, s));
        JSSO_HIP(hipStreamSynchronize(s));
        if (outer == 0) ff = rr;   // u = 0: r = f on the free DOFs
        relmax = 0.0;
        conv = 0;
        for (int k = 0; k < nsys; ++k) {
            const double rel = ff[k] > 0.0 ? std::sqrt(std::max(rr[k], 0.0) / ff[k]) : 0.0;
            relmax = std::max(relmax, rel);
            conv += rel <= tol ? 1 : 0;
        }
        if (conv == nsys || inner_its >= maxit || !std::isfinite(relmax)) break;
        // (a step that gains less than 10%: the attainable accuracy, or no convergence)
        if (outer >= 2 && relmax > 0.9 * prev) break;
        prev = relmax;
        jsso_info ii{};
        p->sym_part_solve = true;
        rc = solve_batched(p, r, du, nrhs, slot, std::max(1e-3, tol), maxit - inner_its, &ii, s);
        p->sym_part_solve = false;
        if (rc < 0) return rc;
        inner_its += ii.iterations;
        if ((rc = launch_axpy((int64_t)ndof * nrhs, du, d_u, s))) return rc;
    }
    if (h_info) {
        *h_info = jsso_info{};
        h_info->iterations = inner_its;
        h_info->replacements = outer;
        h_info->converged = conv;
        h_info->n_systems = nsys;
        h_info->rel_residual = relmax;
    }
    return conv == nsys ? JSSO_OK : JSSO_NOT_CONVERGED;
}

int jsso_solve(jsso_plan* plan, const double* d_f, double* d_u, int32_t nrhs, int32_t precision, double tol,
               int32_t maxit, jsso_info* h_info, void* stream) {
    if (!plan || !d_f || !d_u || nrhs < 1 || maxit < 1 || !(tol > 0.0)) {
        set_error("invalid solve arguments");
        return JSSO_EINVAL;
    }
    Plan* p = P(plan);
    if (!p->assembled) {
        set_error("jsso_solve before jsso_assemble");
        return JSSO_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    if (precision < JSSO_PREC_FP64 || precision > JSSO_PREC_MIXED_V64) {
        set_error("unknown precision %d", precision);
        return JSSO_EINVAL;
    }
    int rc = ensure_device(p);
    if (rc) return rc;
    int32_t status = 0;
    if ((rc = assembly_status(p, s, &status))) return rc;
    if (status & STATUS_NONSYM) return solve_nonsymmetric(p, d_f, d_u, nrhs, precision, tol, maxit, h_info, false, s,
                                                           status);
    switch (precision) {
        case JSSO_PREC_FP64: return solve_batched(p, d_f, d_u, nrhs, PS_F64, tol, maxit, h_info, s);
        case JSSO_PREC_MIXED: return solve_batched(p, d_f, d_u, nrhs, PS_F32, tol, maxit, h_info, s);
        case JSSO_PREC_MIXED_V64: return solve_batched(p, d_f, d_u, nrhs, PS_F32_V64, tol, maxit, h_info, s);
        case JSSO_PREC_PLAIN: return solve_plain(p, d_f, d_u, nrhs, tol, maxit, h_info, s);
        default: set_error("unknown precision %d", precision); return JSSO_EINVAL;
    }
}

int jsso_solve_transpose(jsso_plan* plan, const double* d_f, double* d_u, int32_t nrhs, int32_t precision,
                         double tol, int32_t maxit, jsso_info* h_info, void* stream) {
    if (!plan || !d_f || !d_u || nrhs < 1 || maxit < 1 || !(tol > 0.0)) {
        set_error("invalid solve arguments");
        return JSSO_EINVAL;
    }
    Plan* p = P(plan);
    if (!p->assembled) {
        set_error("jsso_solve_transpose before jsso_assemble");
        return JSSO_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    int rc = ensure_device(p);
    if (rc) return rc;
    int32_t status = 0;
    if ((rc = assembly_status(p, s, &status))) return rc;
    if (status & STATUS_NONSYM) return solve_nonsymmetric(p, d_f, d_u, nrhs, precision, tol, maxit, h_info, true, s,
                                                           status);
    return jsso_solve(plan, d_f, d_u, nrhs, precision, tol, maxit, h_info, stream);
}

int jsso_plan_nonsymmetric_route(jsso_plan* plan, int32_t mode, int32_t* h_mode) {
    if (!plan || mode < -1 || mode > 2) {
        set_error("invalid non-symmetric route %d (-1 query, 0 auto, 1 dense LU, 2 defect correction)", mode);
        return JSSO_EINVAL;
    }
    Plan* p = P(plan);
    if (mode >= 0) p->nonsym_route = mode;
    if (h_mode) *h_mode = p->nonsym_route;
    return JSSO_OK;
}

int jsso_adjoint_grad(jsso_plan* plan, const double* d_crds, const double* d_prop_beam, const double* d_prop_quad,
                      const double* d_u, const double* d_lam, double lam_scale, int32_t nrhs, double* d_dcrds,
                      double* d_dprop_beam, double* d_dprop_quad, void* stream) {
    if (!plan || !d_crds || !d_u || nrhs < 1) {
        set_error("invalid adjoint arguments");
        return JSSO_EINVAL;
    }
    Plan* p = P(plan);
    if ((p->n_beam > 0 && !d_prop_beam) || (p->n_quad > 0 && !d_prop_quad)) {
        set_error("missing element properties");
        return JSSO_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    // the adjoint reads the plan's connectivity and node adjacency on the device
    int rc0 = ensure_device(p);
    if (rc0) return rc0;
    // gradient scratch: per element and coordinate direction, per rhs
    const size_t need = ((size_t)p->n_beam * BDIR + (size_t)p->n_quad * QDIR) * nrhs * sizeof(double);
    if (p->ws.size() <= WS_ADJ_SLOT) p->ws.resize(WS_ADJ_SLOT + 1);
    DBuf& g = p->ws[WS_ADJ_SLOT];
    if (g.bytes < need) {
        if (g.p) JSSO_HIP(hipFree(g.p));
        g.p = nullptr;
        g.bytes = 0;
        hipError_t e = hipMalloc(&g.p, need);
        JSSO_REQUIRE(e == hipSuccess, JSSO_ENOMEM, "gradient scratch hipMalloc(%zu): %s", need, hipGetErrorString(e));
        g.bytes = need;
    }
    std::pair<hipEvent_t, hipEvent_t>* ev = nullptr;
    timing_begin_slot(s, ev, JSSO_TIMING_ADJOINT);
    int rc = launch_adjoint(p, d_crds, d_prop_beam, d_prop_quad, d_u, d_lam, lam_scale, nrhs, (double*)g.p,
                            d_dprop_beam, d_dprop_quad, s);
    if (rc) return rc;
    rc = launch_gather_grad(p, (const double*)g.p, nrhs, d_dcrds, s);
    timing_end(s, ev);
    return rc;
}

int jsso_adjoint_grad_noalloc(jsso_plan* plan, const double* d_crds, const double* d_prop_beam,
                              const double* d_prop_quad, const double* d_u, const double* d_lam, double lam_scale,
                              int32_t nrhs, double* d_dcrds, double* d_dprop_beam, double* d_dprop_quad,
                              void* stream) {
    if (!plan || !d_crds || !d_u || nrhs < 1) {
        set_error("invalid adjoint arguments");
        return JSSO_EINVAL;
    }
    Plan* p = P(plan);
    if ((p->n_beam > 0 && !d_prop_beam) || (p->n_quad > 0 && !d_prop_quad)) {
        set_error("missing element properties");
        return JSSO_EINVAL;
    }
    // uploading the plan allocates: it must have happened before (jsso_assemble does it)
    JSSO_REQUIRE(p->on_device, JSSO_EINVAL,
                 "jsso_adjoint_grad_noalloc: the plan is not on the device yet (call jsso_assemble first)");
    double* gbuf = nullptr;
    size_t cap = 0;
    int rc = noalloc_scratch(&gbuf, &cap);
    if (rc) return rc;
    const size_t need = ((size_t)p->n_beam * BDIR + (size_t)p->n_quad * QDIR) * nrhs * sizeof(double);
    JSSO_REQUIRE(need <= cap, JSSO_EINVAL,
                 "jsso_adjoint_grad_noalloc: needs %zu bytes of scratch ((12 per beam + 17 per quad) x %d load "
                 "cases x 8), the fixed scratch holds %zu; use jsso_adjoint_grad",
                 need, (int)nrhs, cap);
    hipStream_t s = (hipStream_t)stream;
    rc = launch_adjoint(p, d_crds, d_prop_beam, d_prop_quad, d_u, d_lam, lam_scale, nrhs, gbuf, d_dprop_beam,
                        d_dprop_quad, s);
    if (rc) return rc;
    return launch_gather_grad(p, gbuf, nrhs, d_dcrds, s);
}

int jsso_spmv(jsso_plan* plan, const double* d_x, double* d_y, void* stream) {
    if (!plan || !d_x || !d_y) {
        set_error("NULL argument");
        return JSSO_EINVAL;
    }
    Plan* p = P(plan);
    if (!p->assembled) {
        set_error("jsso_spmv before jsso_assemble");
        return JSSO_EINVAL;
    }
    int rc = launch_full_blocks(p, (hipStream_t)stream);
    if (rc) return rc;
    return launch_spmv_full(p, p->d_K, d_x, d_y, (hipStream_t)stream);
}

int jsso_spmv_k(jsso_plan* plan, const double* d_K, const double* d_x, double* d_y, void* stream) {
    if (!plan || !d_x || !d_y) {
        set_error("NULL argument");
        return JSSO_EINVAL;
    }
    Plan* p = P(plan);
    int rc = ensure_device(p);
    if (rc) return rc;
    if (!d_K) {
        if (!p->assembled) {
            set_error("jsso_spmv_k without K values before jsso_assemble");
            return JSSO_EINVAL;
        }
        if ((rc = launch_full_blocks(p, (hipStream_t)stream))) return rc;
        d_K = p->d_K;
    }
    return launch_spmv_full(p, d_K, d_x, d_y, (hipStream_t)stream);
}

int jsso_plan_symmetric(jsso_plan* plan, int32_t* h_symmetric, void* stream) {
    if (!plan || !h_symmetric) {
        set_error("NULL argument");
        return JSSO_EINVAL;
    }
    Plan* p = P(plan);
    if (!p->assembled) {
        set_error("jsso_plan_symmetric before jsso_assemble");
        return JSSO_EINVAL;
    }
    int32_t status = 0;
    int rc = assembly_status(p, (hipStream_t)stream, &status);
    if (rc) return rc;
    *h_symmetric = (status & 2) ? 0 : 1;
    return JSSO_OK;
}

int jsso_dense_solve(jsso_plan* plan, const double* d_K, const double* d_f, double* d_u, int32_t nrhs,
                     int32_t transpose, void* stream) {
    if (!plan || !d_f || !d_u || nrhs < 1 || nrhs > 64) {
        set_error("invalid dense-solve arguments (nrhs 1..64)");
        return JSSO_EINVAL;
    }
    Plan* p = P(plan);
    int rc = ensure_device(p);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    if (!d_K) {
        if (!p->assembled) {
            set_error("jsso_dense_solve without K values before jsso_assemble");
            return JSSO_EINVAL;
        }
        if ((rc = launch_full_blocks(p, s))) return rc;
        d_K = p->d_K;
    }
    int32_t status = 0;
    if (p->assembled && (rc = assembly_status(p, s, &status))) return rc;
    return dense_solve(p, d_K, d_f, d_u, nrhs, (status & 2) != 0, transpose != 0, s);
}

int jsso_csr_spmv(int32_t n_rows, const int32_t* d_rowptr, const int32_t* d_colind, const double* d_vals,
                  const double* d_x, const double* d_c, double* d_y, void* stream) {
    if (n_rows < 0 || (n_rows > 0 && (!d_rowptr || !d_colind || !d_vals || !d_x || !d_y))) {
        set_error("invalid CSR product arguments");
        return JSSO_EINVAL;
    }
    return launch_csr_spmv(n_rows, d_rowptr, d_colind, d_vals, d_x, d_c, d_y, (hipStream_t)stream);
}

int jsso_timing_enable(int32_t every) { return timing_enable(every); }

int jsso_timing_read(double* h_avg_ms, int64_t* h_samples, int32_t reset) {
    return timing_read(JSSO_TIMING_SOLVE, h_avg_ms, h_samples, reset);
}

int jsso_timing_read_slot(int32_t slot, double* h_avg_ms, int64_t* h_samples, int32_t reset) {
    return timing_read(slot, h_avg_ms, h_samples, reset);
}

int jsso_compliance(jsso_plan* plan, const double* d_f, const double* d_u, int32_t nrhs, double scale, double* d_C,
                    void* stream) {
    if (!plan || !d_f || !d_u || !d_C || nrhs < 1) {
        set_error("invalid compliance arguments");
        return JSSO_EINVAL;
    }
    Plan* p = P(plan);
    int rc = ensure_device(p);
    if (rc) return rc;
    return launch_compliance(p, d_f, d_u, nrhs, scale, d_C, (hipStream_t)stream);
}

int jsso_plan_set_preconditioner(jsso_plan* plan, int32_t pc) {
    if (!plan || (pc != JSSO_PC_BLOCK_JACOBI && pc != JSSO_PC_AMG)) {
        set_error("invalid preconditioner (JSSO_PC_BLOCK_JACOBI or JSSO_PC_AMG)");
        return JSSO_EINVAL;
    }
    P(plan)->precond = pc;
    return JSSO_OK;
}
int jsso_plan_amg_reuse(jsso_plan* plan, int32_t mode, int64_t* h_counts) {
    if (!plan || mode < -1 || mode > 2) {
        set_error("invalid AMG reuse mode %d (-1 query, 0 rebuild, 1 keep, 2 rebuild once then keep)", mode);
        return JSSO_EINVAL;
    }
    Plan* p = P(plan);
    if (mode >= 0) {
        p->amg_reuse = mode;
        if (mode == 2) p->amg_valid = false;   // rebuild at the next solve even without a new assembly
    }
    if (h_counts) {
        h_counts[0] = p->amg_reuse;
        h_counts[1] = p->amg_full;
        h_counts[2] = p->amg_light;
    }
    return JSSO_OK;
}

int jsso_plan_amg_levels(jsso_plan* plan, int32_t* h_nodes, int32_t cap, int32_t* h_levels) {
    if (!plan || cap < 0 || (cap > 0 && !h_nodes)) {
        set_error("invalid arguments");
        return JSSO_EINVAL;
    }
    return amg_levels(P(plan), h_nodes, cap, h_levels);
}
int jsso_plan_element_tuning(jsso_plan* plan, int32_t set_epw, int32_t* h_epw) {
    if (!plan || !(set_epw < 0 || set_epw == 0 || set_epw == 4 || set_epw == 8 || set_epw == 16)) {
        set_error("invalid element tuning arguments (epw -1, 0, 4, 8 or 16)");
        return JSSO_EINVAL;
    }
    Plan* p = P(plan);
    if (set_epw >= 0) p->ke_epw = set_epw;
    if (h_epw) *h_epw = p->ke_epw;
    return JSSO_OK;
}

int jsso_plan_cg_tuning(jsso_plan* plan, int32_t precision, int32_t nrhs, int32_t set_wps, int32_t* h_wps) {
    if (!plan || (precision != JSSO_PREC_FP64 && precision != JSSO_PREC_MIXED && precision != JSSO_PREC_MIXED_V64) ||
        nrhs < 1 || nrhs > 8 ||
        !(set_wps < 0 || set_wps == 0 || set_wps == 1 || set_wps == 2 || set_wps == 4)) {
        set_error("invalid tuning arguments (precision FP64 or MIXED, nrhs 1..8, wps -1, 0, 1, 2 or 4)");
        return JSSO_EINVAL;
    }
    const int ps = precision == JSSO_PREC_MIXED ? PS_F32 : precision == JSSO_PREC_FP64 ? PS_F64 : PS_F32_V64;
    Plan* p = P(plan);
    int32_t& slot = p->precond == JSSO_PC_AMG ? p->amg_wps[ps][nrhs] : p->cg_wps[ps][nrhs];
    if (set_wps >= 0) slot = set_wps;
    if (h_wps) *h_wps = slot;
    return JSSO_OK;
}

}  // extern "C"
