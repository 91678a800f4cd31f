// C ABI of the derivative and filter operators (include/mc3d.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>

#include "capi_common.h"
#include "lineop_api.h"
#include "lineop_launch.h"
#include "lineop_plan.h"
#include "mc3d.h"

struct mc3d_lineop {
    mc3d::Plan plan;
    mc3d::OpDev dev{};
    int cls = 0, kind = 0, dir = 0, bc = 0;
    int dims[3] = {0, 0, 0};  // as the kernels see the field: dims[dir] is the line length
    int layout = MC3D_XPENCIL;  // caller's axis order (a rotation of x, y, z)
    int nbatch = 1;         // members stacked along z
    void *dmem = nullptr;
    int device = 0;
    bool fast = false;      // plain applications use the fast kernels
    std::string kname;      // kernel of plain applications
    // transformed-periodic applications run the LDS-tiled kernel, which takes
    // 16- and 32-row chunks only: a 16-row twin when the plain plan has others
    mc3d_lineop *tpb = nullptr;
};

namespace {

int upload(mc3d_lineop *op) {
    const mc3d::Plan &p = op->plan;
    const size_t nr = p.rows.size(), nd = p.drows.size(), nl = p.lu.size(), ns = p.sp.size(),
                 ng = p.G.size();
    const size_t bytes = sizeof(double) * (nr + nd + nl + ns + ng + 2) + sizeof(int) * 2 * p.P + 256;
    if (hipGetDevice(&op->device) != hipSuccess) return mc3d::fail(MC3D_EHIP, "hipGetDevice failed");
    if (hipMalloc(&op->dmem, bytes) != hipSuccess) return mc3d::fail(MC3D_ENOMEM, "hipMalloc of operator tables failed");
    char *base = static_cast<char *>(op->dmem);
    size_t off = 0;
    auto put = [&](const void *src, size_t n) -> void * {
        void *dst = base + off;
        if (n && hipMemcpy(dst, src, n, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
        off += (n + 15) & ~size_t(15);
        return dst;
    };
    double dummy = 0.0;
    mc3d::OpDev &d = op->dev;
    d.rows = static_cast<const double *>(put(nr ? p.rows.data() : &dummy, sizeof(double) * (nr ? nr : 1)));
    d.drows = static_cast<const double *>(put(nd ? p.drows.data() : &dummy, sizeof(double) * (nd ? nd : 1)));
    d.lu = static_cast<const double *>(put(p.lu.data(), sizeof(double) * nl));
    d.sp = static_cast<const double *>(put(p.sp.data(), sizeof(double) * ns));
    d.G = static_cast<const double *>(put(p.G.data(), sizeof(double) * ng));
    d.cstart = static_cast<const int *>(put(p.cstart.data(), sizeof(int) * p.P));
    d.clen = static_cast<const int *>(put(p.clen.data(), sizeof(int) * p.P));
    if (!d.rows || !d.drows || !d.lu || !d.sp || !d.G || !d.cstart || !d.clen)
        return mc3d::fail(MC3D_EHIP, "upload of operator tables failed");
    return MC3D_OK;
}

// The caller's field is x-fastest over the axes rotated by `layout`
// (MC3D_XPENCIL: x, y, z; MC3D_YPENCIL: y, z, x; MC3D_ZPENCIL: z, x, y from
// fastest to slowest, src/Macros.hpp:16-18).  The kernels see it as an
// x-fastest field of dims vd, in which physical direction dir is direction
// vdir and the boundary types follow the same rotation.
int rotate(int layout, int dir, const int dims[3], const int bc3[3], int vd[3], int vb[3], int *vdir) {
    if (layout < MC3D_XPENCIL || layout > MC3D_ZPENCIL) return mc3d::fail(MC3D_EINVAL, "unknown layout");
    if (dir < 0 || dir > 2) return mc3d::fail(MC3D_EINVAL, "dir must be 0, 1 or 2");
    if (!dims) return mc3d::fail(MC3D_EINVAL, "null dims");
    for (int m = 0; m < 3; ++m) {
        vd[m] = dims[(m + layout) % 3];
        if (vb) vb[m] = bc3[(m + layout) % 3];
    }
    *vdir = (dir - layout + 3) % 3;
    return MC3D_OK;
}

int create(int cls, int kind, int dir, int bc, const int bc3[3], const int dims[3], double dd,
           double alphaF, int nbatch, const mc3d::LineTune &tune, mc3d_lineop **out) {
    if (!out) return mc3d::fail(MC3D_EINVAL, "null output handle");
    *out = nullptr;
    if (!dims || dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0)
        return mc3d::fail(MC3D_EINVAL, "dims must be positive");
    if (dir < 0 || dir > 2) return mc3d::fail(MC3D_EINVAL, "dir must be 0, 1 or 2");
    if (bc != MC3D_PERIODIC && bc != MC3D_DIRICHLET) return mc3d::fail(MC3D_EINVAL, "bad boundary type");
    if (cls == mc3d::kDeriv && !(dd > 0.0)) return mc3d::fail(MC3D_EINVAL, "dd must be positive");
    if (nbatch < 1) return mc3d::fail(MC3D_EINVAL, "nbatch must be positive");
    const long npts = (long)dims[0] * dims[1] * dims[2] * nbatch;
    if (npts > (long)2147483647) return mc3d::fail(MC3D_EINVAL, "field larger than 2^31 points");
    if (tune.mm != 0 && tune.mm != 16 && tune.mm != 32) return mc3d::fail(MC3D_EINVAL, "chunk rows must be 16 or 32");
    if (tune.family < 0 || tune.family > 2) return mc3d::fail(MC3D_EINVAL, "bad kernel family");
    if (tune.lines < 0 || (tune.lines & (tune.lines - 1)) != 0)
        return mc3d::fail(MC3D_EINVAL, "lines per workgroup must be a power of two");
    int dev = 0, lds_limit = 65536;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&lds_limit, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
        return mc3d::fail(MC3D_EHIP, "cannot query the device");
    auto *op = new mc3d_lineop();
    std::string err;
    int rc = 0;
    // Chunk rows to try, in order, until one gives a fast kernel: the tuned
    // value alone; else for x-lines that split into 16-row chunks 16 first
    // (contiguous lines are staged whole in LDS, which bounds the waves per
    // CU; 16-row chunks double them, and measured faster for every x operator:
    // the Penta10 derivative of 256-point lines inside the LES step 3.3 ->
    // 2.9 ms), then the plan's own choice (32 rows or an even 8..28-row split),
    // which also serves lines too long for 16-row chunks in one workgroup.
    // When none is fast, 16-row chunks for the LDS-tiled kernel.
    int cand[2] = {tune.mm, -1};
    if (!tune.mm && dir == 0 && dims[0] % 16 == 0) cand[0] = 16, cand[1] = 0;
    auto plan_with = [&](int mm) -> int {
        mc3d::PlanRequest rq{cls, kind, bc, dims[dir], dd, alphaF, mm};
        op->plan = mc3d::Plan();
        int r = mc3d::build_plan(rq, op->plan, err);
        if (r) return r;
        const mc3d::Plan &p = op->plan;
        mc3d::OpDev &d = op->dev;
        d = mc3d::OpDev{};
        d.N = p.N; d.dir = dir; d.W = p.W; d.H = p.H; d.P = p.P; d.L = p.L; d.NI = p.NI;
        d.ext = p.ext; d.MM = p.MM; d.periodic = p.periodic; d.nbl = p.nbl; d.nbr = p.nbr;
        d.int_off = p.int_off; d.int_w = (int)p.int_c.size(); d.gb = p.gb;
        d.cshort = p.cshort; d.lshort = p.lshort;
        d.same_interior = p.same_interior ? 1 : 0;
        for (int k = 0; k < d.int_w; ++k) d.int_c[k] = p.int_c[k];
        d.dd = p.dd; d.divide = p.divide ? 1 : 0;
        // 1/dd is exact when dd is a power of two; then r * (1/dd) == r / dd bitwise
        int ex = 0;
        d.inv_dd = (p.divide && std::frexp(p.dd, &ex) == 0.5) ? 1.0 / p.dd : 0.0;
        d.passA = d.passB = 0;
        if (cls == mc3d::kFilter && bc3) {
            const int da = dir == 0 ? 1 : 0, db = dir == 2 ? 1 : 2;
            d.passA = bc3[da] == MC3D_DIRICHLET;
            d.passB = bc3[db] == MC3D_DIRICHLET;
            d.NA = dims[da];
            d.NB = dims[db];
        }
        const bool rows_ok = p.nbl + p.nbr == 0 || p.dense_ok;
        // the strided fast kernels address a field through a buffer descriptor:
        // unsigned 32-bit byte offsets and record count, so fields below 4 GiB
        // (536 M points: a 640^3 field and up to about 812^3)
        const bool offsets_ok = npts * 8 < 4294967296L;
        d.phases = 1;
        d.phases = mc3d::lineop_fast_phases(d, (size_t)lds_limit);
        op->fast = (p.uniform || p.cshort >= 0) && rows_ok && offsets_ok &&
                   mc3d::lineop_fast_supported(d, (size_t)lds_limit);
        if (tune.family == 2) op->fast = false;
        return MC3D_OK;
    };
    for (int mm : cand) {
        if (mm < 0) break;
        rc = plan_with(mm);
        if (rc) {
            delete op;
            return mc3d::fail(rc, err);
        }
        if (op->fast) break;
    }
    if (!op->fast && !tune.mm && op->plan.MM != 16) {
        rc = plan_with(16);
        if (rc) {
            delete op;
            return mc3d::fail(rc, err);
        }
    }
    if (!op->fast && op->plan.MM != 16 && op->plan.MM != 32) {  // the LDS-tiled kernel's chunk sizes
        delete op;
        return mc3d::fail(MC3D_EINVAL, "no kernel for this line operator and chunking");
    }
    const mc3d::Plan &p = op->plan;
    if (tune.family == 1 && !op->fast) {
        delete op;
        return mc3d::fail(MC3D_EINVAL, "the fast kernels do not take this line operator and chunking");
    }
    if (nbatch > 1 && (op->dev.passA || op->dev.passB)) {
        delete op;
        return mc3d::fail(MC3D_EINVAL, "stacked members need periodic lines in the other directions");
    }
    if (tune.lines) {
        op->dev.L = op->plan.L = tune.lines;
        if (tune.lines * p.P > 256) {
            delete op;
            return mc3d::fail(MC3D_EINVAL, "lines per workgroup times chunks exceeds 256 lanes");
        }
    }
    if (mc3d::lineop_lds_bytes(op->dev) > 65536) {
        delete op;
        return mc3d::fail(MC3D_EINVAL, "line too long for the LDS tile");
    }
    op->cls = cls; op->kind = kind; op->dir = dir; op->bc = bc;
    op->nbatch = nbatch;
    std::memcpy(op->dims, dims, sizeof(op->dims));
    const std::string per = p.periodic ? "true" : "false";
    const bool pair = mc3d::lineop_fast_pair(op->dev);
    if (op->fast)
        op->kname = std::string(dir != 0 ? (pair ? "lineop_strided_pair<" : p.P > 16 ? "lineop_strided_multi<" : "lineop_strided<")
                                         : pair ? "lineop_contig_pair<"
                                         : op->dev.phases > 1 ? "lineop_contig_ph<" : "lineop_contig<") +
                    std::to_string(p.W) +
                    ", " + std::to_string(p.H) + ", " + std::to_string(p.MM) + ", " + per +
                    (p.cshort >= 0 ? ", ragged>" : ">");
    else
        op->kname = "lineop_kernel<" + std::to_string(p.W) + ", " + std::to_string(dir) + ", " +
                    std::to_string(bc == MC3D_PERIODIC ? 0 : 1) + ", " + std::to_string(p.MM) + ">";
    rc = upload(op);
    if (rc) {
        if (op->dmem) (void)hipFree(op->dmem);
        delete op;
        return rc;
    }
    if (cls == mc3d::kDeriv && bc == MC3D_PERIODIC && p.MM != 16 && p.MM != 32) {
        mc3d::LineTune t{};
        t.mm = 16;
        t.family = 2;
        rc = create(cls, kind, dir, bc, bc3, dims, dd, alphaF, nbatch, t, &op->tpb);
        if (rc) {
            mc3d_lineop_destroy(op);
            return rc;
        }
    }
    *out = op;
    return MC3D_OK;
}

int apply(const mc3d_lineop *op, const double *const *in, double *const *out, int nfields,
          int mode, const double *const wrap[6], int accumulate, void *stream) {
    if (!op) return mc3d::fail(MC3D_EINVAL, "null operator");
    if (nfields < 1 || nfields > mc3d::kMaxFields) return mc3d::fail(MC3D_EINVAL, "nfields must be 1..16");
    if (!in || !out) return mc3d::fail(MC3D_EINVAL, "null field list");
    mc3d::LaunchArgs a{};
    for (int f = 0; f < nfields; ++f) {
        if (!in[f] || !out[f]) return mc3d::fail(MC3D_EINVAL, "null field pointer");
        a.in[f] = in[f];
        a.out[f] = out[f];
    }
    for (int k = 0; k < 6; ++k) a.wrap[k] = wrap ? wrap[k] : nullptr;
    if (mode == mc3d::kModeTPB && op->nbatch > 1)
        return mc3d::fail(MC3D_EINVAL, "transformed-periodic derivatives take one member");
    a.Nx = op->dims[0]; a.Ny = op->dims[1]; a.Nz = op->dims[2];
    a.nbatch = op->nbatch;
    // z-lines: one member's lines per grid z layer; x- and y-lines: all members'
    a.nlines = (int)((long)a.Nx * a.Ny * a.Nz * (op->dir == 2 ? 1 : op->nbatch) / op->dims[op->dir]);
    a.accumulate = accumulate;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t e = (op->fast && mode != mc3d::kModeTPB) ? mc3d::lineop_launch_fast(op->dev, a, nfields, st)
                                                         : mc3d::lineop_launch(op->dev, a, nfields, mode, st);
    if (e != hipSuccess) return mc3d::fail(MC3D_EHIP, std::string("line operator launch: ") + hipGetErrorString(e));
    return MC3D_OK;
}

}  // namespace

extern "C" {

int mc3d_deriv_create(int scheme, int dir, int bc, const int dims[3], double dd, int layout,
                      mc3d_lineop **out) {
    if (scheme < MC3D_CD2 || scheme > MC3D_PENTA10) return mc3d::fail(MC3D_EINVAL, "unknown derivative scheme");
    int vd[3], vdir = 0;
    int rc = rotate(layout, dir, dims, nullptr, vd, nullptr, &vdir);
    if (rc) return rc;
    rc = create(mc3d::kDeriv, scheme, vdir, bc, nullptr, vd, dd, 0.0, 1, mc3d::LineTune{}, out);
    if (!rc) (*out)->layout = layout;
    return rc;
}

int mc3d_deriv_apply(const mc3d_lineop *op, const double *const *in, double *const *out, int nfields,
                     void *stream) {
    if (op && op->cls != mc3d::kDeriv) return mc3d::fail(MC3D_EINVAL, "operator is not a derivative");
    if (op && nfields >= 1 && in && out)
        for (int f = 0; f < nfields && f < mc3d::kMaxFields; ++f)
            if (in[f] == out[f]) return mc3d::fail(MC3D_EINVAL, "derivative input and output must not alias");
    const int mode = op && op->bc == MC3D_PERIODIC ? mc3d::kModePlainPeriodic : mc3d::kModePlainDirichlet;
    return apply(op, in, out, nfields, mode, nullptr, 0, stream);
}

int mc3d_deriv_apply_tpb(const mc3d_lineop *op, const double *in, double *out,
                         const double *const wrap[6], void *stream) {
    if (!op || op->cls != mc3d::kDeriv) return mc3d::fail(MC3D_EINVAL, "operator is not a derivative");
    if (op->bc != MC3D_PERIODIC)
        return mc3d::fail(MC3D_EINVAL, "transformed-periodic derivative needs a periodic operator");
    if (!wrap) return mc3d::fail(MC3D_EINVAL, "null wrap list");
    for (int k = 3 - op->plan.H; k < 3 + op->plan.H; ++k)
        if (!wrap[k]) return mc3d::fail(MC3D_EINVAL, "missing wrap array for the scheme's half width");
    if (in == out) return mc3d::fail(MC3D_EINVAL, "derivative input and output must not alias");
    return apply(op->tpb ? op->tpb : op, &in, &out, 1, mc3d::kModeTPB, wrap, 0, stream);
}

int mc3d_filter_create(int kind, double alphaF, int dir, const int bc3[3], const int dims[3], int layout,
                       mc3d_lineop **out) {
    if (kind < MC3D_COMPACT8 || kind > MC3D_COMMN5EXP) return mc3d::fail(MC3D_EINVAL, "unknown filter kind");
    if (!bc3) return mc3d::fail(MC3D_EINVAL, "null bc3");
    int vd[3], vb[3], vdir = 0;
    int rc = rotate(layout, dir, dims, bc3, vd, vb, &vdir);
    if (rc) return rc;
    rc = create(mc3d::kFilter, kind, vdir, vb[vdir], vb, vd, 1.0, alphaF, 1, mc3d::LineTune{}, out);
    if (!rc) (*out)->layout = layout;
    return rc;
}

int mc3d_filter_apply(const mc3d_lineop *op, const double *const *in, double *const *out, int nfields,
                      void *stream) {
    if (op && op->cls != mc3d::kFilter) return mc3d::fail(MC3D_EINVAL, "operator is not a filter");
    const int mode = op && op->bc == MC3D_PERIODIC ? mc3d::kModePlainPeriodic : mc3d::kModePlainDirichlet;
    return apply(op, in, out, nfields, mode, nullptr, 0, stream);
}

void mc3d_lineop_destroy(mc3d_lineop *op) {
    if (!op) return;
    mc3d_lineop_destroy(op->tpb);
    if (op->dmem) {
        int cur = 0;
        (void)hipGetDevice(&cur);
        (void)hipSetDevice(op->device);
        (void)hipFree(op->dmem);
        (void)hipSetDevice(cur);
    }
    delete op;
}

int mc3d_lineop_info(const mc3d_lineop *op, int *chunks, int *chunk_rows, int *lines_per_block) {
    if (!op) return mc3d::fail(MC3D_EINVAL, "null operator");
    if (chunks) *chunks = op->plan.P;
    if (chunk_rows) *chunk_rows = op->plan.clen.empty() ? 0 : op->plan.clen[0];
    if (lines_per_block) *lines_per_block = op->fast ? 64 : op->plan.L;
    return MC3D_OK;
}

const char *mc3d_lineop_kernel(const mc3d_lineop *op) { return op ? op->kname.c_str() : nullptr; }

int mc3d_plan_tables(int cls, int kind, int bc, int n, double dd, double alphaF, int mm,
                     int sizes[11], double *lu, double *sp, double *G, double *rows,
                     double *drows, double *interior, int *cstart, int *clen) {
    if (!sizes) return mc3d::fail(MC3D_EINVAL, "null sizes");
    mc3d::PlanRequest rq{cls, kind, bc, n, dd, alphaF, mm};
    mc3d::Plan p;
    std::string err;
    int rc = mc3d::build_plan(rq, p, err);
    if (rc) return mc3d::fail(rc, err);
    const int v[11] = {p.N, p.W, p.H, p.P, p.NI, p.nbl, p.nbr, p.int_off, p.MM, p.dense_ok ? 1 : 0, p.gb};
    std::memcpy(sizes, v, sizeof(v));
    if (lu) std::memcpy(lu, p.lu.data(), sizeof(double) * p.lu.size());
    if (sp) std::memcpy(sp, p.sp.data(), sizeof(double) * p.sp.size());
    if (G) std::memcpy(G, p.G.data(), sizeof(double) * p.G.size());
    if (rows && !p.rows.empty()) std::memcpy(rows, p.rows.data(), sizeof(double) * p.rows.size());
    if (drows && !p.drows.empty()) std::memcpy(drows, p.drows.data(), sizeof(double) * p.drows.size());
    if (interior) {
        for (int k = 0; k < 2 * mc3d::kHMax + 1; ++k)
            interior[k] = k < (int)p.int_c.size() ? p.int_c[k] : 0.0;
    }
    if (cstart) std::memcpy(cstart, p.cstart.data(), sizeof(int) * p.P);
    if (clen) std::memcpy(clen, p.clen.data(), sizeof(int) * p.P);
    return MC3D_OK;
}

}  // extern "C"

// internal accessors used by the solver
namespace mc3d {
const char *lineop_kernel_name(const mc3d_lineop *op) { return op->kname.c_str(); }

int lineop_create(int cls, int kind, int dir, const int bc3[3], const int dims[3], double dd, double alphaF,
                  int nbatch, const LineTune &tune, mc3d_lineop **out) {
    if (!bc3) return fail(MC3D_EINVAL, "null bc3");
    if (dir < 0 || dir > 2) return fail(MC3D_EINVAL, "dir must be 0, 1 or 2");
    if (cls == kDeriv) {
        if (kind < MC3D_CD2 || kind > MC3D_PENTA10) return fail(MC3D_EINVAL, "unknown derivative scheme");
        return create(cls, kind, dir, bc3[dir], nullptr, dims, dd, 0.0, nbatch, tune, out);
    }
    if (kind < MC3D_COMPACT8 || kind > MC3D_COMMN5EXP) return fail(MC3D_EINVAL, "unknown filter kind");
    return create(cls, kind, dir, bc3[dir], bc3, dims, 1.0, alphaF, nbatch, tune, out);
}

LineTune lineop_shape(const mc3d_lineop *op) {
    LineTune t;
    t.mm = op->plan.MM;
    t.family = op->fast ? 1 : 2;
    t.lines = op->fast ? 0 : op->plan.L;
    return t;
}

int lineop_launch_accumulate(const mc3d_lineop *op, const double *const *in, double *const *out,
                             int nfields, int accumulate, void *stream) {
    const int mode = op->bc == MC3D_PERIODIC ? kModePlainPeriodic : kModePlainDirichlet;
    return apply(op, in, out, nfields, mode, nullptr, accumulate, stream);
}
}  // namespace mc3d
