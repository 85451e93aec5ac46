// LETKF analysis kernels for gfx950 (plain form).
//
// Mathematics of the reference (Hunt et al. 2007 in the form of Zeng et al.
// 2016, src/util/letkf_solver.h:235-349), per gridpoint g with local
// observations o in L(g):
//   ybar = mean_k Y, dY = Y - ybar 1^T, dyo = yo - ybar
//   C    = (M-1)/beta I + dY^T Rl^-1 dY          Rl^-1 = R^-1 o localization
//   C    = V diag(d) V^T
//   w    = V D^-1 V^T dY^T Rl^-1 dyo,  W = sqrt(M-1) V D^-1/2 V^T
//   Xa   = xbar 1^T + dX (W + w 1^T)             X = the 9 populations at g
//
// The observation-space perturbations depend only on the observation, not on the
// gridpoint, so dY is formed once per observation (n_obs x M) and every
// gridpoint gathers its rows through a CSR list of local sites; the reference
// instead re-packs a dense window per gridpoint (src/letkf.cu:150-187) and moves
// it between ranks with all-to-alls.
#include "context.hpp"
#include "philox.hpp"

namespace lbmletkf {

namespace {

struct LocParams {
    int nx, ny, kind, stride, offset, n_sx, n_sy, rloc_len;
    double scale, cutoff;
    const int* site_map;  // explicit site lists: site index of each node or -1; NULL for lattices
};

__host__ __device__ inline double gaspari_cohn(double d) {
    // Gaspari & Cohn (1999) eq. 4.10, as written at src/letkf.cu:105-109
    if (d <= 0) return 1.0;
    if (d <= 1) return -d * d * d * d * d / 4. + d * d * d * d / 2. + d * d * d * 5. / 8. - d * d * 5. / 3. + 1;
    if (d <= 2)
        return d * d * d * d * d / 12. - d * d * d * d / 2. + d * d * d * 5. / 8. + d * d * 5. / 3. - d * 5 + 4 - 2. / 3. / d;
    return 0.0;
}

__host__ __device__ inline int floor_div(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }
__host__ __device__ inline int pmod(int a, int n) { int r = a % n; return r < 0 ? r + n : r; }

// Enumerate the local sites of gridpoint (i,j) in gather order and call
// emit(site, weight).  GC_WINDOW follows src/letkf.cu:29-56 (window geometry)
// and :97-113 (R-localization); GAUSS scans the lattice sites inside the cutoff
// in ascending (sy, sx) order.
template <class Emit>
__device__ inline void enumerate_local(const LocParams& L, int i, int j, Emit emit) {
    const int p = L.stride;
    if (L.kind == LBMLETKF_LOC_GC_WINDOW) {
        int n_obs_local, obs_offset, n_obs_x;
        double d_local;
        if (L.rloc_len <= 0) {
            d_local = p - 1;
            n_obs_local = (p == 1) ? 0 : 2;
            obs_offset = (p == 1) ? 0 : 1;
            n_obs_x = (p == 1) ? 1 : 4;
        } else {
            d_local = double(L.rloc_len) * p;
            obs_offset = 0;
            n_obs_local = 2 * L.rloc_len;
            n_obs_x = 2 * n_obs_local + 1;
        }
        const double cutoff = d_local + 1e-6;
        for (int jr = 0; jr < n_obs_x; ++jr) {
            for (int ir = 0; ir < n_obs_x; ++ir) {
                const int io = (i / p + ir - (n_obs_local - obs_offset)) * p;
                const int jo = (j / p + jr - (n_obs_local - obs_offset)) * p;
                const int di = abs(i - io), dj = abs(j - jo);
                const int iop = pmod(io, L.nx), jop = pmod(jo, L.ny);
                const int site = iop / p + (jop / p) * L.n_sx;
                double w = 1.0;
                if (p > 1) w = gaspari_cohn(sqrt(double(di) * di + double(dj) * dj) / cutoff);
                if (di > n_obs_local * p || dj > n_obs_local * p) w = 0.0;
                emit(site, float(w));
            }
        }
    } else if (L.site_map) {
        // explicit sites: scan the (2R+1)^2 window row by row, ascending (dy, dx)
        const int R = int(floor(L.cutoff));
        const double R2 = L.cutoff * L.cutoff;
        const double inv2L2 = 1.0 / (2.0 * L.scale * L.scale);
        for (int dy = -R; dy <= R; ++dy) {
            const int row = pmod(j + dy, L.ny) * L.nx;
            for (int dx = -R; dx <= R; ++dx) {
                const double d2 = double(dx * dx + dy * dy);
                if (d2 > R2) continue;
                const int site = L.site_map[row + pmod(i + dx, L.nx)];
                if (site >= 0) emit(site, float(exp(-d2 * inv2L2)));
            }
        }
    } else {
        const int R = int(floor(L.cutoff));
        const double R2 = L.cutoff * L.cutoff;
        const double inv2L2 = 1.0 / (2.0 * L.scale * L.scale);
        const int sx0 = -floor_div(-(i - L.offset - R), p), sx1 = floor_div(i - L.offset + R, p);
        const int sy0 = -floor_div(-(j - L.offset - R), p), sy1 = floor_div(j - L.offset + R, p);
        for (int sy = sy0; sy <= sy1; ++sy) {
            const int dy = j - (sy * p + L.offset);
            for (int sx = sx0; sx <= sx1; ++sx) {
                const int dx = i - (sx * p + L.offset);
                const double d2 = double(dx * dx + dy * dy);
                if (d2 <= R2) {
                    const int site = pmod(sx, L.n_sx) + pmod(sy, L.n_sy) * L.n_sx;
                    emit(site, float(exp(-d2 * inv2L2)));
                }
            }
        }
    }
}

__global__ void k_local_count(LocParams L, int* __restrict__ count) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= L.nx * L.ny) return;
    int n = 0;
    enumerate_local(L, g % L.nx, g / L.nx, [&](int, float) { ++n; });
    count[g] = n;
}

__global__ void k_local_fill(LocParams L, const int* __restrict__ ptr, int* __restrict__ site,
                             float* __restrict__ wgt) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= L.nx * L.ny) return;
    int e = ptr[g];
    enumerate_local(L, g % L.nx, g / L.nx, [&](int s, float w) {
        site[e] = s;
        wgt[e] = w;
        ++e;
    });
}

// H(x): moments of each member at each observed site (src/letkf.cu:150-187),
// hx[(s*n_vars + v)*n_members + k].  Evaluated in FP64 from the FP32
// populations: the density is 1 + O(Ma^2), so its ensemble perturbation would
// keep only a few bits in FP32.
__global__ void k_obs_moments(const float* __restrict__ f, const int* __restrict__ site_node,
                              int n_sites, int member0, int n_members, long long member_stride,
                              long long nodes, double c, int n_vars, int v0, int v1, int v2,
                              double* __restrict__ hx) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n_sites * n_members) return;
    const int s = int(t / n_members), k = int(t % n_members);
    const float* fk = f + (long long)(member0 + k) * member_stride + site_node[s];
    double fc[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) fc[q] = fk[q * nodes];
    double r = 0.0;
#pragma unroll
    for (int q = 0; q < Q; ++q) r += fc[q];
    const double u = c * ((fc[2] + fc[5] + fc[8]) - (fc[0] + fc[3] + fc[6])) / r;
    const double v = c * ((fc[6] + fc[7] + fc[8]) - (fc[0] + fc[1] + fc[2])) / r;
    const double val[3] = {u, v, r};
    const int vid[3] = {v0, v1, v2};
    for (int a = 0; a < n_vars; ++a) hx[((long long)s * n_vars + a) * n_members + k] = val[vid[a]];
}

__device__ inline double wave_sum(double x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    return x;
}

// ybar, dY and the innovation per observation; one wavefront per observation
// (src/util/letkf_solver.h:239-243).  A non-finite observation sets status bit 1.
__global__ __launch_bounds__(64) void k_obs_prepare(const double* __restrict__ hx,
                                                    const float* __restrict__ yo, int n_obs, int M,
                                                    float* __restrict__ dy,
                                                    float* __restrict__ dyo,
                                                    int* __restrict__ status) {
    const int o = blockIdx.x;
    const int k = threadIdx.x;
    const double x = (k < M) ? hx[(long long)o * M + k] : 0.0;
    const double ybar = wave_sum(x) / double(M);
    if (k < M) dy[(long long)o * M + k] = float(x - ybar);
    if (k == 0) {
        const float y = yo[o];
        if (!isfinite(y)) atomicOr(status, 2);
        dyo[o] = float(double(y) - ybar);
    }
}

// y = H(x) + sigma_var * N(0,1), Philox keyed by (seed, counter) -- the OBSERVE
// step of src/io.cu:160-246 restricted to the observed sites.
__global__ void k_observe(const double* __restrict__ hx, float* __restrict__ yo, int n_obs,
                          int n_vars, int v0, int v1, int v2, double sig_u, double sig_rho,
                          uint64_t seed, uint64_t counter, int noise) {
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n_obs) return;
    double y = hx[o];
    if (noise) {
        double z[4];
        philox_normal4(seed, uint32_t(counter), uint32_t(counter >> 32), uint32_t(o / 4), STREAM_OBS, z);
        const int vid[3] = {v0, v1, v2};
        const double sig = vid[o % n_vars] == 2 ? sig_rho : sig_u;
        y += sig * z[o % 4];
    }
    yo[o] = float(y);
}

// ----------------------------------------------------------------------------
// plain Gram matrix: C = (M-1)/beta I + dY^T Rl^-1 dY and b = dY^T Rl^-1 dyo in
// FP32 on the vector ALUs, one 256-thread workgroup per gridpoint, 4x4 register
// blocks of C per thread, observation rows staged through LDS in chunks.
// Replaces the two strided-batched GEMMs of src/util/letkf_solver.h:251-254
// (the reference multiplies by a dense diagonal rR first).
// ----------------------------------------------------------------------------
constexpr int GRAM_NT = 256, GRAM_CH = 32, GRAM_MP = 64;

__global__ __launch_bounds__(GRAM_NT) void k_gram_fp32(
    const int* __restrict__ loc_ptr, const int* __restrict__ loc_site,
    const float* __restrict__ loc_w, const float* __restrict__ dy, const float* __restrict__ dyo,
    const float* __restrict__ coef, int n_vars, int M, float diag, float* __restrict__ gram,
    float* __restrict__ bvec) {
    __shared__ float sD[GRAM_CH][GRAM_MP];
    __shared__ float sW[GRAM_CH], sE[GRAM_CH];
    const int g = blockIdx.x;
    const int t = threadIdx.x;
    const int NB = (M + 3) / 4;
    const bool active = t < NB * NB;
    const int kb = t % NB, lb = t / NB;
    float acc[4][4] = {};
    float bacc[4] = {};
    const int e0 = loc_ptr[g], e1 = loc_ptr[g + 1];
    const int total = (e1 - e0) * n_vars;
    for (int base = 0; base < total; base += GRAM_CH) {
        for (int idx = t; idx < GRAM_CH * GRAM_MP; idx += GRAM_NT) {
            const int r = idx / GRAM_MP, k = idx % GRAM_MP;
            const int fi = base + r;
            float val = 0.f;
            if (fi < total) {
                const int e = e0 + fi / n_vars, va = fi % n_vars;
                const long long o = (long long)loc_site[e] * n_vars + va;
                if (k < M) val = dy[o * M + k];
                if (k == 0) {
                    sW[r] = loc_w[e] * coef[va];
                    sE[r] = dyo[o];
                }
            } else if (k == 0) {
                sW[r] = 0.f;
                sE[r] = 0.f;
            }
            sD[r][k] = val;
        }
        __syncthreads();
        if (active) {
            for (int r = 0; r < GRAM_CH; ++r) {
                const float w = sW[r];
                float a[4], bb[4];
#pragma unroll
                for (int x = 0; x < 4; ++x) {
                    a[x] = sD[r][4 * kb + x];
                    bb[x] = sD[r][4 * lb + x];
                }
#pragma unroll
                for (int x = 0; x < 4; ++x) {
                    const float wa = w * a[x];
#pragma unroll
                    for (int y = 0; y < 4; ++y) acc[x][y] += wa * bb[y];
                }
                if (lb == 0) {
                    const float we = w * sE[r];
#pragma unroll
                    for (int x = 0; x < 4; ++x) bacc[x] += we * a[x];
                }
            }
        }
        __syncthreads();
    }
    if (!active) return;
    float* Cg = gram + (long long)g * M * M;
#pragma unroll
    for (int x = 0; x < 4; ++x) {
        const int k = 4 * kb + x;
        if (k >= M) continue;
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            const int l = 4 * lb + y;
            if (l >= M) continue;
            Cg[k * M + l] = acc[x][y] + (k == l ? diag : 0.f);
        }
        if (lb == 0) bvec[(long long)g * M + k] = bacc[x];
    }
}

// ----------------------------------------------------------------------------
// plain solve: one wavefront per gridpoint, FP64 cyclic Jacobi with A and V
// resident in LDS, then the transform and the in-place state update.
//
// Ordering: N-1 rounds per sweep; round m pairs every index p with p XOR m, so
// each round is a perfect matching and a sweep visits every pair exactly once
// (N a power of two; smaller ensembles are padded with decoupled unit rows).
// The reference calls cusolverDn{S,D}syevjBatched / EigenG-Batched for this
// step (src/util/cu_matrix_eigensolver_wrapper.hpp:63-76,223-238).
// ----------------------------------------------------------------------------
struct SolveParams {
    int M, nodes;
    long long member_stride;
    double tol2;
    int max_sweeps;
    int dump_g0, dump_n;
    unsigned long long* sweeps;  // optional: Jacobi sweeps summed over gridpoints
};

__device__ inline void xor_pair(int k, int m, int& p, int& q) {
    const int hb = 31 - __clz(m);
    p = ((k >> hb) << (hb + 1)) | (k & ((1 << hb) - 1));
    q = p ^ m;
}

template <int N>
__global__ __launch_bounds__(64) void k_solve_fp64(const float* __restrict__ gram,
                                                   const float* __restrict__ bvec,
                                                   float* __restrict__ f, SolveParams sp,
                                                   double* __restrict__ dump,
                                                   int* __restrict__ status) {
    constexpr int NP = N / 2;
    constexpr int LD = N + 1;
    __shared__ double A[N * LD];
    __shared__ double Vt[N * LD];  // Vt[i*LD + r] = V[r][i]: column i of V is contiguous
    __shared__ double rc[NP], rs[NP], rt[NP];
    __shared__ double hv[N], sq[N], wv[N], bv[N];
    __shared__ double dX[Q * N];
    __shared__ double xbar[Q];

    const int g = blockIdx.x;
    const int lane = threadIdx.x;
    const int M = sp.M;
    const float* Cg = gram + (long long)g * M * M;
    for (int idx = lane; idx < N * N; idx += 64) {
        const int i = idx / N, j = idx % N;
        A[i * LD + j] = (i < M && j < M) ? double(Cg[i * M + j]) : (i == j ? 1.0 : 0.0);
        Vt[i * LD + j] = (i == j) ? 1.0 : 0.0;
    }
    for (int i = lane; i < N; i += 64) bv[i] = (i < M) ? double(bvec[(long long)g * M + i]) : 0.0;
    const bool dumping = (g >= sp.dump_g0) && (g < sp.dump_g0 + sp.dump_n);
    double* dg = dumping ? dump + (long long)(g - sp.dump_g0) * dump_record(M) : nullptr;
    if (dumping)  // C as the eigensolve takes it
        for (int idx = lane; idx < M * M; idx += 64) dg[M * M + 2 * M + idx] = double(Cg[idx]);
    __syncthreads();

    int sweep = 0;
    for (; sweep < sp.max_sweeps; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int idx = lane; idx < N * N; idx += 64) {
            const int i = idx / N, j = idx % N;
            const double a = A[i * LD + j];
            tot += a * a;
            if (i != j) off += a * a;
        }
        off = wave_sum(off);
        tot = wave_sum(tot);
        if (off <= sp.tol2 * tot) break;
        for (int m = 1; m < N; ++m) {
            if (lane < NP) {
                int p, q;
                xor_pair(lane, m, p, q);
                const double app = A[p * LD + p], aqq = A[q * LD + q], apq = A[p * LD + q];
                double c = 1.0, s = 0.0, t = 0.0;
                if (apq != 0.0) {
                    const double theta = (aqq - app) / (2.0 * apq);
                    t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                    c = 1.0 / sqrt(t * t + 1.0);
                    s = t * c;
                }
                rc[lane] = c;
                rs[lane] = s;
                rt[lane] = t;
            }
            __syncthreads();
            for (int blk = lane; blk < NP * NP; blk += 64) {
                const int I = blk / NP, J = blk % NP;
                if (I > J) continue;
                int p, q, r, s2;
                xor_pair(I, m, p, q);
                if (I == J) {
                    const double apq = A[p * LD + q], t = rt[I];
                    A[p * LD + p] -= t * apq;
                    A[q * LD + q] += t * apq;
                    A[p * LD + q] = 0.0;
                    A[q * LD + p] = 0.0;
                } else {
                    xor_pair(J, m, r, s2);
                    const double cI = rc[I], sI = rs[I], cJ = rc[J], sJ = rs[J];
                    const double b11 = A[p * LD + r], b12 = A[p * LD + s2];
                    const double b21 = A[q * LD + r], b22 = A[q * LD + s2];
                    const double x11 = cI * b11 - sI * b21, x12 = cI * b12 - sI * b22;
                    const double x21 = sI * b11 + cI * b21, x22 = sI * b12 + cI * b22;
                    const double y11 = cJ * x11 - sJ * x12, y12 = sJ * x11 + cJ * x12;
                    const double y21 = cJ * x21 - sJ * x22, y22 = sJ * x21 + cJ * x22;
                    A[p * LD + r] = y11; A[p * LD + s2] = y12;
                    A[q * LD + r] = y21; A[q * LD + s2] = y22;
                    A[r * LD + p] = y11; A[s2 * LD + p] = y12;
                    A[r * LD + q] = y21; A[s2 * LD + q] = y22;
                }
            }
            if (lane < N) {
                for (int k = 0; k < NP; ++k) {
                    int p, q;
                    xor_pair(k, m, p, q);
                    const double c = rc[k], s = rs[k];
                    const double vp = Vt[p * LD + lane], vq = Vt[q * LD + lane];
                    Vt[p * LD + lane] = c * vp - s * vq;
                    Vt[q * LD + lane] = s * vp + c * vq;
                }
            }
            __syncthreads();
        }
    }

    if (sp.sweeps && lane == 0) atomicAdd(sp.sweeps, (unsigned long long)sweep);

    // transform: w = V D^-1 V^T b, W = sqrt(M-1) V D^-1/2 V^T  (letkf_solver.h:288-306)
    const double sqm1 = sqrt(double(M - 1));
    if (lane < N) {
        const double d = A[lane * LD + lane];
        double gsum = 0.0;
        for (int k = 0; k < N; ++k) gsum += Vt[lane * LD + k] * bv[k];
        hv[lane] = gsum / d;
        sq[lane] = sqm1 / sqrt(d);
    }
    __syncthreads();
    double dval = 0.0;
    if (lane < N) {
        dval = A[lane * LD + lane];
        double ws = 0.0;
        for (int i = 0; i < N; ++i) ws += Vt[i * LD + lane] * hv[i];
        wv[lane] = ws;
    }
    __syncthreads();
    // column `lane` of W into A (A is no longer needed): T = W + w 1^T
    if (dumping && lane < M)  // V[k][lane] = Vt[lane][k], the eigenvectors in the solver's order
        for (int k = 0; k < M; ++k) dg[2 * M * M + 2 * M + k * M + lane] = Vt[lane * LD + k];
    if (lane < N) {
        double ucol[N];
#pragma unroll
        for (int i = 0; i < N; ++i) ucol[i] = Vt[i * LD + lane] * sq[i];
        for (int k = 0; k < N; ++k) {
            double acc = 0.0;
#pragma unroll
            for (int i = 0; i < N; ++i) acc += Vt[i * LD + k] * ucol[i];
            if (dumping && k < M && lane < M) dg[k * M + lane] = acc;
            // A[k][lane] holds T[k][lane]; row k of A is only read below
            A[k * LD + lane] = acc + wv[k];
        }
    }
    if (dumping && lane < M) {
        dg[M * M + lane] = wv[lane];
        dg[M * M + M + lane] = dval;
    }
    // Xa = xbar + dX T (letkf_solver.h:313-316), written back into f at g
    float xk[Q];
    const int member = lane < M ? lane : 0;  // lanes >= M hold no member
    const long long base = (long long)member * sp.member_stride + g;
#pragma unroll
    for (int q = 0; q < Q; ++q) xk[q] = 0.f;
    if (lane < M) {
#pragma unroll
        for (int q = 0; q < Q; ++q) xk[q] = f[base + (long long)q * sp.nodes];
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const double sum = wave_sum(double(xk[q]));
        if (lane == 0) xbar[q] = sum / double(M);
    }
    __syncthreads();
    if (lane < N) {
#pragma unroll
        for (int q = 0; q < Q; ++q) dX[q * N + lane] = (lane < M) ? double(xk[q]) - xbar[q] : 0.0;
    }
    __syncthreads();
    if (lane < M) {
        bool bad = false;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            double acc = xbar[q];
            for (int k = 0; k < M; ++k) acc += dX[q * N + k] * A[k * LD + lane];
            const float xa = float(acc);
            bad |= !isfinite(xa);
            f[base + (long long)q * sp.nodes] = xa;
        }
        if (bad) atomicOr(status, 1);
    }
}

LocParams make_loc(const Ctx& c) {
    LocParams L;
    L.nx = c.nx;
    L.ny = c.ny;
    L.kind = c.cfg.loc_kind;
    L.stride = c.obs.stride;
    L.offset = c.obs.offset;
    L.n_sx = c.obs.n_sx;
    L.n_sy = c.obs.n_sy;
    L.rloc_len = c.cfg.loc_rloc_len;
    L.scale = c.cfg.loc_scale;
    L.cutoff = c.cfg.loc_cutoff;
    L.site_map = c.obs.custom ? c.d_site_map : nullptr;
    return L;
}

}  // namespace

hipError_t launch_local_count(Ctx& c, int* d_count) {
    const int nt = 256;
    k_local_count<<<int((c.nodes + nt - 1) / nt), nt, 0, c.stream>>>(make_loc(c), d_count);
    return hipGetLastError();
}

hipError_t launch_local_fill(Ctx& c) {
    const int nt = 256;
    k_local_fill<<<int((c.nodes + nt - 1) / nt), nt, 0, c.stream>>>(make_loc(c), c.d_loc_ptr,
                                                                      c.d_loc_site, c.d_loc_w);
    return hipGetLastError();
}

hipError_t launch_obs_moments(Ctx& c, const float* f, double* hx, int member0, int n_members) {
    const int nt = 256;
    const long long n = (long long)c.obs.n_sites * n_members;
    if (n == 0) return hipSuccess;  // a network without sites
    k_obs_moments<<<int((n + nt - 1) / nt), nt, 0, c.stream>>>(
        f, c.d_site_node, c.obs.n_sites, member0, n_members, c.lp.member_stride, c.nodes,
        c.cfg.c, c.obs.n_vars, c.obs.var_id[0], c.obs.var_id[1], c.obs.var_id[2], hx);
    return hipGetLastError();
}

hipError_t launch_obs_prepare(Ctx& c) {
    const int n_obs = c.obs.n_sites * c.obs.n_vars;
    if (n_obs == 0) return hipSuccess;
    k_obs_prepare<<<n_obs, 64, 0, c.stream>>>(c.d_hx, c.d_yo, n_obs, c.M, c.d_dy, c.d_dyo,
                                              c.d_status);
    return hipGetLastError();
}

hipError_t launch_observe(Ctx& c, const double* hx_member, float* yo, uint64_t seed,
                          uint64_t counter, int noise) {
    const int n_obs = c.obs.n_sites * c.obs.n_vars;
    if (n_obs == 0) return hipSuccess;
    const int nt = 256;
    k_observe<<<(n_obs + nt - 1) / nt, nt, 0, c.stream>>>(
        hx_member, yo, n_obs, c.obs.n_vars, c.obs.var_id[0], c.obs.var_id[1], c.obs.var_id[2],
        c.cfg.obs_sigma_u, c.cfg.obs_sigma_rho, seed, counter, noise);
    return hipGetLastError();
}

hipError_t launch_gram_fp32(Ctx& c) {
    if (!c.d_gram || !c.d_bvec || !c.d_loc_ptr || !c.d_dy || !c.d_dyo) return hipErrorInvalidValue;
    const float diag = float(double(c.M - 1) / c.cfg.inflation);
    k_gram_fp32<<<int(c.nodes), GRAM_NT, 0, c.stream>>>(c.d_loc_ptr, c.d_loc_site, c.d_loc_w,
                                                        c.d_dy, c.d_dyo, c.d_obs_coef,
                                                        c.obs.n_vars, c.M, diag, c.d_gram,
                                                        c.d_bvec);
    return hipGetLastError();
}

hipError_t launch_solve_fp64(Ctx& c, float* f, unsigned long long* sweeps) {
    if (!c.d_gram || !c.d_bvec || !f || (c.dump_n > 0 && !c.d_dump)) return hipErrorInvalidValue;
    SolveParams sp;
    sp.sweeps = sweeps;
    sp.M = c.M;
    sp.nodes = int(c.nodes);
    sp.member_stride = c.lp.member_stride;
    sp.tol2 = c.cfg.jacobi_tol * c.cfg.jacobi_tol;
    sp.max_sweeps = c.cfg.jacobi_max_sweeps;
    sp.dump_g0 = c.dump_g0;
    sp.dump_n = c.dump_n;
    int N = 2;
    while (N < c.M) N *= 2;
    const dim3 grid(unsigned(c.nodes)), block(64);
    switch (N) {
        case 2: k_solve_fp64<2><<<grid, block, 0, c.stream>>>(c.d_gram, c.d_bvec, f, sp, c.d_dump, c.d_status); break;
        case 4: k_solve_fp64<4><<<grid, block, 0, c.stream>>>(c.d_gram, c.d_bvec, f, sp, c.d_dump, c.d_status); break;
        case 8: k_solve_fp64<8><<<grid, block, 0, c.stream>>>(c.d_gram, c.d_bvec, f, sp, c.d_dump, c.d_status); break;
        case 16: k_solve_fp64<16><<<grid, block, 0, c.stream>>>(c.d_gram, c.d_bvec, f, sp, c.d_dump, c.d_status); break;
        case 32: k_solve_fp64<32><<<grid, block, 0, c.stream>>>(c.d_gram, c.d_bvec, f, sp, c.d_dump, c.d_status); break;
        case 64: k_solve_fp64<64><<<grid, block, 0, c.stream>>>(c.d_gram, c.d_bvec, f, sp, c.d_dump, c.d_status); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace lbmletkf
