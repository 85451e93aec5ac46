// D2Q9 BGK + Smagorinsky forecast kernels for gfx950.
//
// Physics follows the reference step update() (src/lbm.h:16-29):
//   forcing (src/data.cu:305-350) -> pull streaming + moments (src/lbm.cu:67-103)
//   -> Smagorinsky nu_t from central differences of the streamed velocity
//   (src/lbm.cu:24-65) -> BGK relaxation with tau = 0.5 + 3(nu+nu_t)/(c^2 dt)
//   and the first-order force term (src/lbm.cu:105-160, src/lbm.h:31-70).
//
// Two forms are built:
//   * lbm_fused: one launch advances every member by one step.  A 64x16 tile of
//     one member is streamed into registers, its streamed velocity (plus a one-node
//     halo) goes to LDS for the strain stencil, and the collided populations are
//     written once: 36 B read + 36 B written per node update.  A workgroup may
//     take several tiles of a row band in turn (one per workgroup is fastest at
//     256^2 x 64: the tiles of one workgroup do not overlap their loads).
//   * the plain form keeps the reference's three passes (stream+macro, les,
//     collide) and is launched member by member, as one reference rank would.
#include <map>
#include <type_traits>
#include <mutex>

#include "context.hpp"
#include "lbm_node.hpp"
#include "philox.hpp"

namespace lbmletkf {

namespace {

using namespace lbmnode;

// ----------------------------------------------------------------------------
// fused, batched step
// ----------------------------------------------------------------------------
constexpr int TX = 64, TY = 16, NT = 256;
constexpr int ROWS_PER_PASS = NT / TX;        // 4
constexpr int NODES_PER_THREAD = TY / ROWS_PER_PASS;  // 4
constexpr int HALO = 2 * (TX + 2) + 2 * TY;   // 164

// Tile t (member-major order of 64x16 tiles) of one step.
template <bool FORCE>
__device__ __forceinline__ void lbm_tile(int t, const float* __restrict__ fin, float* __restrict__ fout,
                                         const float* __restrict__ force, const LbmParams& p, int tiles_x,
                                         int tiles_per_member) {
    const int nx = p.nx, ny = p.ny;
    const long long nodes = p.nodes;
    __shared__ float su[TY + 2][TX + 2];
    __shared__ float sv[TY + 2][TX + 2];
    const int lx = threadIdx.x % TX;
    const int ly0 = threadIdx.x / TX;
    const int member = t / tiles_per_member;
    const int tt = t - member * tiles_per_member;
    const int x0 = (tt % tiles_x) * TX;
    const int y0 = (tt / tiles_x) * TY;
    const float* f = fin + (long long)member * p.member_stride;
    float* fo = fout + (long long)member * p.member_stride;
    const int i = x0 + lx;
    const int im = wrap(i - 1, nx), ip = wrap(i + 1, nx);

    __syncthreads();  // the previous tile's stencil reads of su/sv are done
    float fc[NODES_PER_THREAD][Q];
    float rr[NODES_PER_THREAD], uu[NODES_PER_THREAD], vv[NODES_PER_THREAD];
#pragma unroll
    for (int r = 0; r < NODES_PER_THREAD; ++r) {
        const int ly = ly0 + r * ROWS_PER_PASS;
        const int j = y0 + ly;
        const long long r0 = (long long)j * nx;
        const long long rm = (long long)wrap(j - 1, ny) * nx;
        const long long rp = (long long)wrap(j + 1, ny) * nx;
        pull9(f, nodes, im, i, ip, rm, r0, rp, fc[r]);
        moments(fc[r], p.c, rr[r], uu[r], vv[r]);
        su[ly + 1][lx + 1] = uu[r];
        sv[ly + 1][lx + 1] = vv[r];
    }
    // halo ring of streamed velocities
    if (threadIdx.x < HALO) {
        int h = threadIdx.x, hx, hy;
        if (h < TX + 2) { hx = h - 1; hy = -1; }
        else if (h < 2 * (TX + 2)) { hx = h - (TX + 2) - 1; hy = TY; }
        else if (h < 2 * (TX + 2) + TY) { hx = -1; hy = h - 2 * (TX + 2); }
        else { hx = TX; hy = h - 2 * (TX + 2) - TY; }
        const int gi = wrap(x0 + hx, nx), gj = wrap(y0 + hy, ny);
        const int gim = wrap(gi - 1, nx), gip = wrap(gi + 1, nx);
        const long long r0 = (long long)gj * nx;
        const long long rm = (long long)wrap(gj - 1, ny) * nx;
        const long long rp = (long long)wrap(gj + 1, ny) * nx;
        float fh[Q], rh, uh, vh;
        pull9(f, nodes, gim, gi, gip, rm, r0, rp, fh);
        moments(fh, p.c, rh, uh, vh);
        su[hy + 1][hx + 1] = uh;
        sv[hy + 1][hx + 1] = vh;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < NODES_PER_THREAD; ++r) {
        const int ly = ly0 + r * ROWS_PER_PASS;
        const int j = y0 + ly;
        const float nut = smagorinsky(su[ly + 1][lx + 2], su[ly + 1][lx], su[ly + 2][lx + 1],
                                      su[ly][lx + 1], sv[ly + 1][lx + 2], sv[ly + 1][lx],
                                      sv[ly + 2][lx + 1], sv[ly][lx + 1], p);
        const long long node = (long long)j * nx + i;
        float Fx = 0.f, Fy = 0.f;
        if (FORCE && p.has_force) {
            Fx = force[node];
            Fy = force[nodes + node];
        }
        collide9<FORCE>(fc[r], rr[r], uu[r], vv[r], nut, Fx, Fy, p);
#pragma unroll
        for (int q = 0; q < Q; ++q) fo[q * nodes + node] = fc[r][q];
    }
}

template <bool FORCE>
__global__ __launch_bounds__(NT) void k_lbm_fused(const float* __restrict__ fin,
                                                  float* __restrict__ fout,
                                                  const float* __restrict__ force, LbmParams p,
                                                  int tiles_x, int tiles_per_member, int tiles_per_block) {
    // Each workgroup advances tiles_per_block consecutive tiles (a row band of one
    // member), so the grid is one resident generation with no tail.  XCD-aware
    // order: workgroups are dealt round-robin to the 8 XCDs, so each XCD gets a
    // contiguous run of bands (neighbouring bands share halo rows in its L2).
    const int nb = gridDim.x;
    const int b = blockIdx.x;
    const int bb = (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
    for (int k = 0; k < tiles_per_block; ++k)
        lbm_tile<FORCE>(bb * tiles_per_block + k, fin, fout, force, p, tiles_x, tiles_per_member);
}

// Full-row strips: a workgroup of nx threads (one per column, nx a multiple of
// 64 up to 1024) advances rows y0 .. y0+H-1 of one member, one row at a time.
// The streamed velocities of the last rows sit in a 4-row LDS ring, so a row's
// Smagorinsky stencil needs no halo in x (the row is whole and periodic) and only
// the two rows y0-1 and y0+H are streamed twice (by neighbouring strips).  The
// populations of row k+2 are requested before the barrier of row k.
template <bool FORCE, int H>
__global__ __launch_bounds__(1024) void k_lbm_rows(const float* __restrict__ fin, float* __restrict__ fout,
                                                   const float* __restrict__ force, LbmParams p,
                                                   int strips_per_member) {
    extern __shared__ float ring[];  // [2][4][nx]: u, v of the rows k-1 .. k+2
    const int nb = gridDim.x, b = blockIdx.x;
    const int bb = (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
    const int nx = p.nx, ny = p.ny;
    const long long nodes = p.nodes;
    const int member = bb / strips_per_member;
    const int y0 = (bb - member * strips_per_member) * H;
    const float* f = fin + (long long)member * p.member_stride;
    float* fo = fout + (long long)member * p.member_stride;
    float* ru = ring;
    float* rv = ring + 4 * nx;
    const int i = threadIdx.x;
    const int im = i == 0 ? nx - 1 : i - 1, ip = i == nx - 1 ? 0 : i + 1;
    auto pull = [&](int jrel, float fc[Q]) {
        const int j = wrap(y0 + jrel, ny);
        pull9(f, nodes, im, i, ip, (long long)wrap(j - 1, ny) * nx, (long long)j * nx,
              (long long)wrap(j + 1, ny) * nx, fc);
    };
    auto vel = [&](int jrel, const float fc[Q], float& r, float& u, float& v) {
        moments(fc, p.c, r, u, v);
        ru[((jrel + 4) & 3) * nx + i] = u;
        rv[((jrel + 4) & 3) * nx + i] = v;
    };
    float fa[Q], fb[Q], fn[Q], ra, ua, va, rb, ub, vb, rx, ux, vx;
    pull(-1, fn);
    vel(-1, fn, rx, ux, vx);
    pull(0, fa);
    vel(0, fa, ra, ua, va);
    pull(1, fb);
    for (int k = 0; k < H; ++k) {
        if (k + 2 <= H) pull(k + 2, fn);
        vel(k + 1, fb, rb, ub, vb);
        __syncthreads();  // rows k-1 .. k+1 are in the ring; row k+2 overwrites the slot of k-2
        const int s = ((k + 4) & 3) * nx, sn = ((k + 5) & 3) * nx, ss = ((k + 3) & 3) * nx;
        const float nut = smagorinsky(ru[s + ip], ru[s + im], ru[sn + i], ru[ss + i], rv[s + ip], rv[s + im],
                                      rv[sn + i], rv[ss + i], p);
        const long long node = (long long)(y0 + k) * nx + i;
        float Fx = 0.f, Fy = 0.f;
        if (FORCE && p.has_force) {
            Fx = force[node];
            Fy = force[nodes + node];
        }
        collide9<FORCE>(fa, ra, ua, va, nut, Fx, Fy, p);
#pragma unroll
        for (int q = 0; q < Q; ++q) fo[q * nodes + node] = fa[q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            fa[q] = fb[q];
            fb[q] = fn[q];
        }
        ra = rb;
        ua = ub;
        va = vb;
    }
}

// Two steps per launch (no body force, no friction): a workgroup of nx threads
// (one per column, nx <= 1024) produces rows [y0, y0+H) of one member at t+2 from
// the populations at t; the intermediate step never leaves the chip.  Iteration k
// runs four stages, each on another row and each followed by a barrier:
//   A (row k):   pull from global (t), moments -> step-1 velocity ring u1/v1
//   B (row k-1): Smagorinsky from u1/v1, collide -> populations at t+1 -> ring p1
//   C (row k-2): pull from p1, moments -> step-2 velocity ring u2/v2
//   D (row k-3): Smagorinsky from u2/v2, collide -> global (t+2)
// Stage A covers strip rows -3 .. H+2 (it reads rows -4 .. H+3; the outer rows
// are also read by the neighbouring strips, mostly from L2), B rows -2 .. H+1,
// C rows -1 .. H and D rows 0 .. H-1.  Every ring holds 3 rows (a stage reads rows
// r-1, r, r+1 of its ring); a slot is rewritten only after the stage that read it
// last has passed two barriers.  The ring p1
// holds the six populations that stream across columns (ex != 0); the three with
// ex = 0 (q = 1, 4, 7) stream within the column, so each thread keeps its own in
// registers (one and two rows back).  Each node runs the arithmetic of the
// single-step kernels (pull9, moments, smagorinsky, collide9), so two of these
// steps equal two single steps bit for bit.  HBM traffic: 36 B read + 36 B
// written per node per launch, i.e. per two steps.
constexpr int P1Q = 6;  // populations in the ring: q = 0, 2, 3, 5, 6, 8
__device__ __forceinline__ constexpr int p1_slot(int q) { return (q / 3) * 2 + (q % 3 == 2 ? 1 : 0); }
template <int H>
__global__ __launch_bounds__(1024) void k_lbm_rows2(const float* __restrict__ fin, float* __restrict__ fout,
                                                   LbmParams p, int strips_per_member) {
    extern __shared__ float lds2[];  // u1, v1, u2, v2: [3][nx] each; p1: [3][P1Q][nx]
    const int nx = p.nx, ny = p.ny;
    const int nodes = int(p.nodes);  // offsets within one member fit in 32 bits
    float* u1 = lds2;
    float* v1 = u1 + 3 * nx;
    float* u2 = v1 + 3 * nx;
    float* v2 = u2 + 3 * nx;
    float* p1 = v2 + 3 * nx;
    const int nb = gridDim.x, b = blockIdx.x;
    const int bb = (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
    const int member = bb / strips_per_member;
    const int y0 = (bb - member * strips_per_member) * H;
    const float* f = fin + (long long)member * p.member_stride;
    float* fo = fout + (long long)member * p.member_stride;
    const int i = threadIdx.x;
    const int im = i == 0 ? nx - 1 : i - 1, ip = i == nx - 1 ? 0 : i + 1;
    auto sl = [](int r) { return (r + 9) % 3; };  // ring slot of strip row r >= -9
    // unsigned 32-bit offsets from the member's base: the loads and stores address
    // as uniform base + per-lane offset
    const unsigned un = unsigned(nodes), ui = unsigned(i), uim = unsigned(im), uip = unsigned(ip);
    auto pull_g = [&](int r, float (&fc)[Q]) {
        const int j = wrap(y0 + r, ny);
        const unsigned rm = unsigned(wrap(j - 1, ny) * nx), r0 = unsigned(j * nx), rp = unsigned(wrap(j + 1, ny) * nx);
        fc[0] = f[0 * un + rp + uip];
        fc[1] = f[1 * un + rp + ui];
        fc[2] = f[2 * un + rp + uim];
        fc[3] = f[3 * un + r0 + uip];
        fc[4] = f[4 * un + r0 + ui];
        fc[5] = f[5 * un + r0 + uim];
        fc[6] = f[6 * un + rm + uip];
        fc[7] = f[7 * un + rm + ui];
        fc[8] = f[8 * un + rm + uim];
    };
    float q4m1 = 0.f, q7m1 = 0.f, q7m2 = 0.f;  // f1[4] of row k-2, f1[7] of rows k-2, k-3
    float fP[Q];
    // One iteration.  A writes (fa, ma) for row k; B reads the previous A row (fb,
    // mb); C writes (fc, mc) for row k-2; D reads the previous C row (fd, md).  The
    // loop runs iterations in pairs with the two register sets swapped, so no row
    // is copied from one set to the other.
    auto iter = [&](int k, float (&fa)[Q], float (&ma)[3], float (&fb)[Q], float (&mb)[3], float (&fc)[Q],
                    float (&mc)[3], float (&fd)[Q], float (&md)[3]) {
        // A: row k (the pull was issued one iteration earlier)
#pragma unroll
        for (int q = 0; q < Q; ++q) fa[q] = fP[q];
        if (k < H + 2) pull_g(k + 1, fP);
        moments(fa, p.c, ma[0], ma[1], ma[2]);
        u1[sl(k) * nx + i] = ma[1];
        v1[sl(k) * nx + i] = ma[2];
        __syncthreads();
        // B: row k-1, the populations at t+1
        if (k >= -1) {
            const int s = sl(k - 1) * nx, sn = sl(k) * nx, ss = sl(k - 2) * nx;
            const float nut = smagorinsky(u1[s + ip], u1[s + im], u1[sn + i], u1[ss + i], v1[s + ip],
                                          v1[s + im], v1[sn + i], v1[ss + i], p);
            collide9<false>(fb, mb[0], mb[1], mb[2], nut, 0.f, 0.f, p);
            float* pr = p1 + sl(k - 1) * P1Q * nx;
#pragma unroll
            for (int q = 0; q < Q; ++q)
                if (q % 3 != 1) pr[p1_slot(q) * nx + i] = fb[q];
        }
        __syncthreads();
        // C: row k-2, pulled from the ring (row above = k-1, below = k-3) and, for
        // q = 1, 4, 7, from this thread's own rows
        if (k >= 1) {
            const float* pp = p1 + sl(k - 1) * P1Q * nx;
            const float* p0 = p1 + sl(k - 2) * P1Q * nx;
            const float* pm = p1 + sl(k - 3) * P1Q * nx;
            fc[0] = pp[p1_slot(0) * nx + ip];
            fc[1] = fb[1];
            fc[2] = pp[p1_slot(2) * nx + im];
            fc[3] = p0[p1_slot(3) * nx + ip];
            fc[4] = q4m1;
            fc[5] = p0[p1_slot(5) * nx + im];
            fc[6] = pm[p1_slot(6) * nx + ip];
            fc[7] = q7m2;
            fc[8] = pm[p1_slot(8) * nx + im];
            moments(fc, p.c, mc[0], mc[1], mc[2]);
            u2[sl(k - 2) * nx + i] = mc[1];
            v2[sl(k - 2) * nx + i] = mc[2];
        }
        q7m2 = q7m1;
        q7m1 = fb[7];
        q4m1 = fb[4];
        __syncthreads();
        // D: row k-3, the populations at t+2
        if (k >= 3) {
            const int s = sl(k - 3) * nx, sn = sl(k - 2) * nx, ss = sl(k - 4) * nx;
            const float nut = smagorinsky(u2[s + ip], u2[s + im], u2[sn + i], u2[ss + i], v2[s + ip],
                                          v2[s + im], v2[sn + i], v2[ss + i], p);
            collide9<false>(fd, md[0], md[1], md[2], nut, 0.f, 0.f, p);
            const unsigned node = unsigned((y0 + k - 3) * nx) + ui;
#pragma unroll
            for (int q = 0; q < Q; ++q) fo[unsigned(q) * un + node] = fd[q];
        }
    };
    float fX[Q], fY[Q], fZ[Q], fW[Q], mX[3], mY[3] = {0.f, 0.f, 0.f}, mZ[3], mW[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; q < Q; ++q) fY[q] = fW[q] = 0.f;
    pull_g(-3, fP);
    static_assert((H + 6) % 2 == 0, "iterations come in pairs");
    for (int k = -3; k <= H + 2; k += 2) {
        iter(k, fX, mX, fY, mY, fZ, mZ, fW, mW);
        iter(k + 1, fY, mY, fX, mX, fW, mW, fZ, mZ);
    }
}

// Three steps per launch (no body force, no friction), the two-step pipeline
// extended by two stages: iteration k runs six stages a row apart:
//   A (row k):   pull from global (t), moments -> velocity ring 1
//   B (row k-1): Smagorinsky, collide -> populations at t+1 -> population ring 1
//   C (row k-2): pull from population ring 1, moments -> velocity ring 2
//   D (row k-3): Smagorinsky, collide -> populations at t+2 -> population ring 2
//   E (row k-4): pull from population ring 2, moments -> velocity ring 3
//   F (row k-5): Smagorinsky, collide -> populations at t+3 -> global
// A strip of H rows reads rows -6 .. H+5 and runs step one on rows -4 .. H+3,
// step two on -2 .. H+1 (each step has a halo of two rows: one for the pull, one
// for the central differences of the Smagorinsky viscosity).
// The rings keep each value only as long as other columns read it, so that four
// workgroups of nx = 256 fit one CU (144 B per column, 36 KB):
// * velocity rings: row r in slot r & 1; the collide of row r reads the east and
//   west neighbours of row r there, while its own column's velocities of rows
//   r - 1 and r + 1 stay in the thread's registers;
// * population rings hold the six populations that stream across columns, by
//   when they are read: the pull of row r takes q = 0, 2 from row r + 1 (one slot:
//   written and read in one iteration), q = 3, 5 from row r (two slots) and
//   q = 6, 8 from row r - 1 (three slots); q = 1, 4, 7 (e_x = 0) stream within
//   the column and stay in registers.
// Two barriers per iteration suffice, after B and after D: the only values a
// stage reads from the LDS in the iteration that wrote them are the q = 0, 2
// populations of the row just collided; every other ring slot is read one or two
// iterations after its write and rewritten after a barrier that follows its last
// read.
// Every node runs the per-node functions of the single-step kernels, so three of
// these steps equal three single steps bit for bit.  HBM traffic: 36 B read + 36 B
// written per node per launch, i.e. per three steps.
struct PopRing {
    float* q02;  // [2][nx]     q = 0, 2 of the last collided row
    float* q35;  // [2][2][nx]  q = 3, 5 of rows of parity 0 / 1
    float* q68;  // [3][2][nx]  q = 6, 8 of rows r mod 3
};
// NX > 0: the row length as a compile-time constant (every LDS ring offset then
// becomes an immediate of its instruction); NX = 0 takes p.nx.
// Odd strips are stepped as the lattice reflected in y: logical row r is physical
// row y0 + H - 1 - r, and population q is the physical population with e_y
// negated (0, 1, 2 <-> 6, 7, 8), so those strips walk their rows bottom-up.  The
// per-node functions are symmetric under that reflection, so the bits are those
// of the upward walk; the reflection changes only the uniform (scalar) row and
// population offsets of the loads and stores.  Two neighbouring strips then reach
// their common boundary rows together (at the start of both, or at the end of
// both), and the second read of those halo rows finds them in L2.
template <int H, int NX>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void k_lbm_rows3(
    const float* __restrict__ fin, float* __restrict__ fout, LbmParams p, int strips_per_member) {
    extern __shared__ float lds3[];  // velocity rings [3 stages][u, v][2][nx]; population rings [2][12][nx]
    const int nx = NX > 0 ? NX : p.nx, ny = p.ny;
    const int nodes = int(p.nodes);  // offsets within one member fit in 32 bits
    float* vel = lds3;
    PopRing r1{vel + 12 * nx, vel + 14 * nx, vel + 18 * nx};
    PopRing r2{vel + 24 * nx, vel + 26 * nx, vel + 30 * nx};
    const int nb = gridDim.x, b = blockIdx.x;
    const int bb = (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
    const int member = bb / strips_per_member;
    const int strip = bb - member * strips_per_member;
    const int y0 = strip * H;
    const bool mir = (strip & 1) != 0;  // uniform
    // the member's populations through buffer resources: the byte offset of
    // population q in row j, (q nodes + j nx) 4, is uniform and goes to the scalar
    // offset; the only per-lane offset is the column, 4 i (no vector address
    // arithmetic per load or store).  One member is 36 nodes bytes (< 2^31 for the
    // grids this kernel takes).
    const __amdgpu_buffer_rsrc_t fr = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float*>(fin + (long long)member * p.member_stride), 0, Q * nodes * 4, 0x00020000);
    const __amdgpu_buffer_rsrc_t fw =
        __builtin_amdgcn_make_buffer_rsrc(fout + (long long)member * p.member_stride, 0, Q * nodes * 4, 0x00020000);
    const int i = threadIdx.x;
    const int im = i == 0 ? nx - 1 : i - 1, ip = i == nx - 1 ? 0 : i + 1;
    const int bi = 4 * i, bim = 4 * im, bip = 4 * ip;
    auto ld = [&](int voff, int soff) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(fr, voff, soff, 0));
    };
    // pull of logical row r: fc[q] = f[q](i - ex_q, r - ey_q) of the (reflected) lattice
    auto pull_g = [&](int r, float (&fc)[Q]) {
        const int j = wrap(mir ? y0 + H - 1 - r : y0 + r, ny);
        const int rm = 4 * wrap(j - 1, ny) * nx, r0 = 4 * j * nx, rp = 4 * wrap(j + 1, ny) * nx;
        const int qn = 4 * nodes;
        // physical offsets of logical population 0 (pulled from logical row r + 1) and
        // logical population 6 (from row r - 1); 1, 2 and 7, 8 follow them
        const int os = mir ? 6 * qn + rm : rp, on = mir ? rp : 6 * qn + rm;
        fc[0] = ld(bip, os);
        fc[1] = ld(bi, os + qn);
        fc[2] = ld(bim, os + 2 * qn);
        fc[3] = ld(bip, 3 * qn + r0);
        fc[4] = ld(bi, 4 * qn + r0);
        fc[5] = ld(bim, 5 * qn + r0);
        fc[6] = ld(bip, on);
        fc[7] = ld(bi, on + qn);
        fc[8] = ld(bim, on + 2 * qn);
    };
    // LDS column offsets of this thread and its west and east neighbours, renewed
    // from an opaque copy at every iteration: formed once outside the loop, the
    // ring addresses built on them would stay live across it and spill
    int ci = i, cm = im, cp = ip;
    auto q35_slot = [&](const PopRing& R, int r) { return R.q35 + (r & 1) * 2 * nx; };
    auto q68_slot = [&](const PopRing& R, int r) { return R.q68 + ((r + 12) % 3) * 2 * nx; };
    // pull of row r from a population ring; q = 1, 4, 7 from the thread's own rows
    auto pull_ring = [&](const PopRing& R, int r, float q1, float q4, float q7, float (&fc)[Q]) {
        const float* a = R.q02;               // row r + 1
        const float* c = q35_slot(R, r);      // row r
        const float* d = q68_slot(R, r - 1);  // row r - 1
        fc[0] = a[cp];
        fc[1] = q1;
        fc[2] = a[nx + cm];
        fc[3] = c[cp];
        fc[4] = q4;
        fc[5] = c[nx + cm];
        fc[6] = d[cp];
        fc[7] = q7;
        fc[8] = d[nx + cm];
    };
    auto to_ring = [&](const PopRing& R, int r, const float (&fd)[Q]) {
        R.q02[ci] = fd[0];
        R.q02[nx + ci] = fd[2];
        float* c = q35_slot(R, r);
        c[ci] = fd[3];
        c[nx + ci] = fd[5];
        float* d = q68_slot(R, r);
        d[ci] = fd[6];
        d[nx + ci] = fd[8];
    };
    // velocity ring of stage s (0, 1, 2): u of row r at vel[(s * 4 + (r & 1)) * nx],
    // v two slots further
    auto vel_put = [&](int s, int r, float u, float v) {
        vel[(s * 4 + (r & 1)) * nx + ci] = u;
        vel[(s * 4 + 2 + (r & 1)) * nx + ci] = v;
    };
    // collide of row r: own velocities of rows r+1 (uN, vN) and r-1 (uS, vS) from
    // registers, those of the east and west neighbours of row r from the ring
    auto collide_row = [&](int s, int r, float uN, float vN, float uS, float vS, float (&fd)[Q],
                           const float (&md)[3]) {
        const float* ur = vel + (s * 4 + (r & 1)) * nx;
        const float* vr = vel + (s * 4 + 2 + (r & 1)) * nx;
        const float nut = smagorinsky(ur[cp], ur[cm], uN, uS, vr[cp], vr[cm], vN, vS, p);
        collide9<false>(fd, md[0], md[1], md[2], nut, 0.f, 0.f, p);
    };
    // e_x = 0 populations of the collided rows: q4 of the last row, q7 of the last two
    float a4m1 = 0.f, a7m1 = 0.f, a7m2 = 0.f;  // step one
    float b4m1 = 0.f, b7m1 = 0.f, b7m2 = 0.f;  // step two
    // rows pulled ahead: two rows in flight, the even one in fP0, the odd one in fP1
    float fP0[Q], fP1[Q];
    // One iteration; the stage pairs (A, B), (C, D), (E, F) alternate register sets
    // between consecutive iterations, so no row is copied from one set to the other.
    // A moments stage's set holds, on entry, the moments of its row two iterations
    // back: the south velocities of its collide stage.
    // The parity of k is a compile-time constant of each of the two calls per loop
    // step (the loop starts at an odd k), so that every ring slot chosen by r & 1 is
    // a constant offset.
    auto iter = [&](auto parity, int k, float (&fa)[Q], float (&ma)[3], float (&fb)[Q], float (&mb)[3],
                    float (&fc)[Q], float (&mc)[3], float (&fd)[Q], float (&md)[3], float (&fe)[Q], float (&me)[3],
                    float (&ff)[Q], float (&mf)[3]) {
        k = (k & ~1) | decltype(parity)::value;
        ci = i;
        asm volatile("" : "+v"(ci));
        cm = ci == 0 ? nx - 1 : ci - 1;
        cp = ci == nx - 1 ? 0 : ci + 1;
        // A: row k (the pull was issued two iterations earlier)
        const float uSa = ma[1], vSa = ma[2];
        float (&fP)[Q] = decltype(parity)::value ? fP1 : fP0;
#pragma unroll
        for (int q = 0; q < Q; ++q) fa[q] = fP[q];
        if (k + 2 <= H + 4) pull_g(k + 2, fP);
        moments(fa, p.c, ma[0], ma[1], ma[2]);
        vel_put(0, k, ma[1], ma[2]);
        __builtin_amdgcn_sched_barrier(0);  // stages stay in order (register pressure)
        // B: row k-1 -> populations at t+1
        if (k >= -3) {
            collide_row(0, k - 1, ma[1], ma[2], uSa, vSa, fb, mb);
            to_ring(r1, k - 1, fb);
        }
        __syncthreads();
        // C: row k-2 at t+1
        const float uSc = mc[1], vSc = mc[2];
        if (k >= -1) {
            pull_ring(r1, k - 2, fb[1], a4m1, a7m2, fc);
            moments(fc, p.c, mc[0], mc[1], mc[2]);
            vel_put(1, k - 2, mc[1], mc[2]);
        }
        a7m2 = a7m1;
        a7m1 = fb[7];
        a4m1 = fb[4];
        __builtin_amdgcn_sched_barrier(0);
        // D: row k-3 -> populations at t+2
        if (k >= 1) {
            collide_row(1, k - 3, mc[1], mc[2], uSc, vSc, fd, md);
            to_ring(r2, k - 3, fd);
        }
        __syncthreads();
        // E: row k-4 at t+2
        const float uSe = me[1], vSe = me[2];
        if (k >= 3) {
            pull_ring(r2, k - 4, fd[1], b4m1, b7m2, fe);
            moments(fe, p.c, me[0], me[1], me[2]);
            vel_put(2, k - 4, me[1], me[2]);
        }
        b7m2 = b7m1;
        b7m1 = fd[7];
        b4m1 = fd[4];
        __builtin_amdgcn_sched_barrier(0);
        // F: row k-5 -> populations at t+3, written out
        if (k >= 5) {
            collide_row(2, k - 5, me[1], me[2], uSe, vSe, ff, mf);
            const int row = 4 * (mir ? y0 + H - 1 - (k - 5) : y0 + k - 5) * nx;
            // physical offsets of logical populations 0 and 6 (the reflection swaps them)
            const int o0 = (mir ? 24 : 0) * nodes + row, o6 = (mir ? 0 : 24) * nodes + row;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int off = q < 3 ? o0 + 4 * q * nodes : (q < 6 ? 4 * q * nodes + row : o6 + 4 * (q - 6) * nodes);
                __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, ff[q]), fw, bi, off, 0);
            }
        }
    };
    float fX[Q], fY[Q], fZ[Q], fW[Q], fS[Q], fT[Q];
    float mX[3] = {0.f, 0.f, 0.f}, mY[3] = {0.f, 0.f, 0.f}, mZ[3] = {0.f, 0.f, 0.f}, mW[3] = {0.f, 0.f, 0.f},
          mS[3] = {0.f, 0.f, 0.f}, mT[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; q < Q; ++q) fY[q] = fW[q] = fT[q] = 0.f;
    pull_g(-5, fP1);
    pull_g(-4, fP0);
    static_assert((H + 10) % 2 == 0, "iterations come in pairs");
    for (int k = -5; k <= H + 4; k += 2) {  // k odd
        iter(std::integral_constant<int, 1>{}, k, fX, mX, fY, mY, fZ, mZ, fW, mW, fS, mS, fT, mT);
        iter(std::integral_constant<int, 0>{}, k + 1, fY, mY, fX, mX, fW, mW, fZ, mZ, fT, mT, fS, mS);
    }
}



// ----------------------------------------------------------------------------
// plain form: the reference's three passes for one member
// ----------------------------------------------------------------------------
__global__ void k_stream_macro(const float* __restrict__ f, float* __restrict__ fn,
                               float* __restrict__ ruv, LbmParams p) {
    const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= p.nodes) return;
    const int nx = p.nx, ny = p.ny;
    const int i = int(node % nx), j = int(node / nx);
    float fc[Q];
    pull9(f, p.nodes, wrap(i - 1, nx), i, wrap(i + 1, nx), (long long)wrap(j - 1, ny) * nx,
          (long long)j * nx, (long long)wrap(j + 1, ny) * nx, fc);
#pragma unroll
    for (int q = 0; q < Q; ++q) fn[q * p.nodes + node] = fc[q];
    float r, u, v;
    moments(fc, p.c, r, u, v);
    ruv[node] = r;
    ruv[p.nodes + node] = u;
    ruv[2 * p.nodes + node] = v;
}

__global__ void k_les(const float* __restrict__ ruv, float* __restrict__ nus, LbmParams p) {
    const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= p.nodes) return;
    const int nx = p.nx, ny = p.ny;
    const int i = int(node % nx), j = int(node / nx);
    const float* u = ruv + p.nodes;
    const float* v = ruv + 2 * p.nodes;
    const long long e = (long long)j * nx + wrap(i + 1, nx), w = (long long)j * nx + wrap(i - 1, nx);
    const long long n = (long long)wrap(j + 1, ny) * nx + i, s = (long long)wrap(j - 1, ny) * nx + i;
    nus[node] = smagorinsky(u[e], u[w], u[n], u[s], v[e], v[w], v[n], v[s], p);
}

__global__ void k_collide(float* __restrict__ fn, const float* __restrict__ ruv,
                          const float* __restrict__ nus, const float* __restrict__ force,
                          LbmParams p) {
    const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= p.nodes) return;
    float fc[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) fc[q] = fn[q * p.nodes + node];
    float Fx = 0.f, Fy = 0.f;
    if (p.has_force) {
        Fx = force[node];
        Fy = force[p.nodes + node];
    }
    collide9<true>(fc, ruv[node], ruv[p.nodes + node], ruv[2 * p.nodes + node], nus[node], Fx, Fy, p);
#pragma unroll
    for (int q = 0; q < Q; ++q) fn[q * p.nodes + node] = fc[q];
}

// ----------------------------------------------------------------------------
// stochastic forcing of Watanabe (1997): one field per step, shared by all
// members (the reference seeds every member's generator identically,
// src/data.cu:227-228).  f_x = sum_n amp_n 2 pi ky_n [(r0 kx + r1 ky) sin(th)
// + (r2 kx + r3 ky) cos(th)], f_y likewise with -kx   (src/data.cu:305-349)
// ----------------------------------------------------------------------------
__global__ void k_force_field(const ForceTable* __restrict__ tab, uint64_t seed, float stdev,
                              const long long* __restrict__ d_step, int step_offset,
                              float* __restrict__ force, int nx, int ny) {
    __shared__ float rnd[4][MAX_FORCE_MODES];
    __shared__ int kxs[MAX_FORCE_MODES], kys[MAX_FORCE_MODES];
    __shared__ float amps[MAX_FORCE_MODES];
    const int n = tab->n;
    const long long step = *d_step + step_offset;
    for (int m = threadIdx.x; m < n; m += blockDim.x) {
        double z[4];
        philox_normal4(seed, uint32_t(step), uint32_t(uint64_t(step) >> 32), uint32_t(m),
                       STREAM_FORCE, z);
        for (int a = 0; a < 4; ++a) rnd[a][m] = float(z[a] * double(stdev));
        kxs[m] = int(tab->kx[m]);
        kys[m] = int(tab->ky[m]);
        amps[m] = tab->amp[m];
    }
    __syncthreads();
    const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nodes = (long long)nx * ny;
    if (node >= nodes) return;
    const int i = int(node % nx), j = int(node / nx);
    // phase 2 pi (kx x + ky y), x = (i - nx/2)/nx, reduced exactly in integers
    const long long period = (long long)nx * ny;
    const float two_pi = 6.28318530717958647692f;
    float fx = 0.f, fy = 0.f;
    for (int m = 0; m < n; ++m) {
        const int kx = kxs[m], ky = kys[m];
        long long num = (long long)kx * (i - nx / 2) * ny + (long long)ky * (j - ny / 2) * nx;
        num %= period;
        if (num < 0) num += period;
        float s, c;
        sincospif(2.f * float(double(num) / double(period)), &s, &c);
        const float fek = (rnd[0][m] * kx + rnd[1][m] * ky) * s + (rnd[2][m] * kx + rnd[3][m] * ky) * c;
        fx += amps[m] * two_pi * float(ky) * fek;
        fy += amps[m] * two_pi * float(-kx) * fek;
    }
    force[node] = fx;
    force[nodes + node] = fy;
}

__global__ void k_step_advance(long long* d_step, int n) { *d_step += n; }

// f := feq(rho, u, v) (src/lbm.h:31-39)
__global__ void k_set_feq(float* __restrict__ f, const float* __restrict__ rho,
                          const float* __restrict__ u, const float* __restrict__ v, LbmParams p) {
    const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= p.nodes) return;
    float feq[Q];
    equilibrium9(rho[node], u[node], v[node], p, feq);
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q * p.nodes + node] = feq[q];
}

// feq_q(rho, u, v) of every q by the formula (src/lbm.h:31-39), as the reference's
// nudging and perturbation evaluate it per population
__device__ __forceinline__ float feq_formula(float rho, float u, float v, int q, const LbmParams& p) {
    const int ex = q % 3 - 1, ey = q / 3 - 1;
    const float wq = (ex == 0 && ey == 0) ? W0 : ((ex == 0 || ey == 0) ? W1 : W2);
    const float uc = (u * float(ex) + v * float(ey)) * p.inv_c;
    const float uu = (u * u + v * v) * p.inv_c2;
    return rho * wq * (1.f + 3.f * uc + 4.5f * uc * uc - 1.5f * uu);
}

// Newtonian nudging of every member towards the observed fields
// (src/nudging.cu:40-142 with da_w_itp): the observation lattice (stride xyprune)
// is interpolated bilinearly to each node (directly read when xyprune = 1), and
// f_q := alpha feq_q(rho_o, u_o, v_o) + (1 - alpha) f_q.  skip_east / skip_north
// leave the nodes with i >= nx/2 / j >= ny/2 untouched (DA_QUADRA); without an
// observed density (use_obs_rho = 0) rho_o is the member's own density.
__global__ void k_nudge(float* __restrict__ f, const float* __restrict__ ro, const float* __restrict__ uo,
                        const float* __restrict__ vo, LbmParams p, float alpha, int xyprune, int skip_east,
                        int skip_north, int use_obs_rho) {
    const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= p.nodes) return;
    const int nx = p.nx, ny = p.ny;
    const int i = int(node % nx), j = int(node / nx);
    if ((skip_east && i >= nx / 2) || (skip_north && j >= ny / 2)) return;
    float rob, uob, vob;
    if (xyprune == 1) {
        rob = use_obs_rho ? ro[node] : 0.f;
        uob = uo[node];
        vob = vo[node];
    } else {
        const int iw = (i / xyprune) * xyprune, js = (j / xyprune) * xyprune;
        const int ie = (iw + xyprune) % nx, jn = (js + xyprune) % ny;
        const long long k0 = (long long)js * nx + iw, ke = (long long)js * nx + ie;
        const long long kn = (long long)jn * nx + iw, ken = (long long)jn * nx + ie;
        const double ir = double(float(i - iw) / float(xyprune)), jr = double(float(j - js) / float(xyprune));
        const double c0 = (1. - ir) * (1. - jr), ce = ir * (1. - jr), cn = (1. - ir) * jr, cen = ir * jr;
        uob = float(c0 * uo[k0] + ce * uo[ke] + cn * uo[kn] + cen * uo[ken]);
        vob = float(c0 * vo[k0] + ce * vo[ke] + cn * vo[kn] + cen * vo[ken]);
        rob = use_obs_rho ? float(c0 * ro[k0] + ce * ro[ke] + cn * ro[kn] + cen * ro[ken]) : 0.f;
    }
    for (int m = 0; m < p.n_members; ++m) {
        float* fm = f + (long long)m * p.member_stride;
        float fc[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) fc[q] = fm[q * p.nodes + node];
        float r = rob;
        if (!use_obs_rho) {
            r = 0.f;
#pragma unroll
            for (int q = 0; q < Q; ++q) r += fc[q];
        }
#pragma unroll
        for (int q = 0; q < Q; ++q)
            fm[q * p.nodes + node] = alpha * feq_formula(r, uob, vob, q, p) + (1.f - alpha) * fc[q];
    }
}

// Perturbation of the Lyapunov runs (src/data.cu:426-456): with random velocities
// (ur, vr) per node, f_q := eps feq_q(rho, ur, vr) + (1 - eps) f_q, rho the node's
// density.
__global__ void k_perturb(float* __restrict__ f, const float* __restrict__ ur, const float* __restrict__ vr,
                          LbmParams p, float eps) {
    const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= p.nodes) return;
    float fc[Q];
    float r = 0.f;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        fc[q] = f[q * p.nodes + node];
        r += fc[q];
    }
#pragma unroll
    for (int q = 0; q < Q; ++q)
        f[q * p.nodes + node] = eps * feq_formula(r, ur[node], vr[node], q, p) + (1.f - eps) * fc[q];
}

// rho, u, v of the stored populations (no streaming), src/letkf.cu:336-357
__global__ void k_macro(const float* __restrict__ f, float* __restrict__ rho,
                        float* __restrict__ u, float* __restrict__ v, LbmParams p) {
    const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= p.nodes) return;
    float fc[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) fc[q] = f[q * p.nodes + node];
    float r, uu, vv;
    moments(fc, p.c, r, uu, vv);
    if (rho) rho[node] = r;
    if (u) u[node] = uu;
    if (v) v[node] = vv;
}

// (u, v) of every member at every node: uv[member][0][node] = u, [1] = v
__global__ void k_member_uv(const float* __restrict__ f, float* __restrict__ uv, LbmParams p) {
    const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= p.nodes) return;
    const int m = blockIdx.y;
    const float* fm = f + (long long)m * p.member_stride;
    float fc[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) fc[q] = fm[q * p.nodes + node];
    float r, uu, vv;
    moments(fc, p.c, r, uu, vv);
    uv[(2LL * m) * p.nodes + node] = uu;
    uv[(2LL * m + 1) * p.nodes + node] = vv;
}

// Per node over the members (src/io.cu:90-145 with ENSEMBLE_STAT, and the
// vorticity of src/io.cu:61-86): out[0] = mean u, [1] = mean v, [2] = mean
// vorticity, [3] = its standard deviation, [4], [5] = standard deviations of u
// and v; deviations biased (1/M), from FP64 sums of values and squares.
__global__ void k_ensemble_fields(const float* __restrict__ uv, float* __restrict__ out, int M, float dx,
                                  LbmParams p) {
    const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= p.nodes) return;
    const int nx = p.nx, ny = p.ny;
    const int i = int(node % nx), j = int(node / nx);
    const long long e = (long long)j * nx + (i + 1 == nx ? 0 : i + 1);
    const long long w = (long long)j * nx + (i == 0 ? nx - 1 : i - 1);
    const long long n = (long long)(j + 1 == ny ? 0 : j + 1) * nx + i;
    const long long s = (long long)(j == 0 ? ny - 1 : j - 1) * nx + i;
    double su = 0, sv = 0, su2 = 0, sv2 = 0, sw = 0, sw2 = 0;
    for (int m = 0; m < M; ++m) {
        const float* u = uv + (2LL * m) * p.nodes;
        const float* v = u + p.nodes;
        const float uy = (u[n] - u[s]) / (2.f * dx);
        const float vx = (v[e] - v[w]) / (2.f * dx);
        const float vor = vx - uy;
        const double a = u[node], b = v[node];
        su += a;
        sv += b;
        su2 += a * a;
        sv2 += b * b;
        sw += vor;
        sw2 += double(vor) * vor;
    }
    const double im = 1.0 / M;
    auto sd = [](double ss, double mean) {
        const double var = ss - mean * mean;
        return var > 0 ? sqrt(var) : 0.0;
    };
    const double mu = su * im, mv = sv * im, mw = sw * im;
    out[node] = float(mu);
    out[p.nodes + node] = float(mv);
    out[2 * p.nodes + node] = float(mw);
    out[3 * p.nodes + node] = float(sd(sw2 * im, mw));
    out[4 * p.nodes + node] = float(sd(su2 * im, mu));
    out[5 * p.nodes + node] = float(sd(sv2 * im, mv));
}

// Ensemble mean / biased spread of (u,v) and rho against a truth, summed per
// block in FP64 (ensemble_stat/main.cpp:108-177).  partials[block][4] =
// { sum (ubar-ut)^2+(vbar-vt)^2, sum su^2+sv^2, sum (rbar-rt)^2, sum sr^2 }.
constexpr int STATS_NT = 256;
__global__ __launch_bounds__(STATS_NT) void k_ensemble_stats(
    const float* __restrict__ f, const float* __restrict__ trho, const float* __restrict__ tu,
    const float* __restrict__ tv, double* __restrict__ partials, LbmParams p) {
    __shared__ double red[4][STATS_NT];
    const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double acc[4] = {0, 0, 0, 0};
    if (node < p.nodes) {
        double su = 0, sv = 0, sr = 0, suu = 0, svv = 0, srr = 0;
        const int M = p.n_members;
        for (int k = 0; k < M; ++k) {
            const float* fk = f + (long long)k * p.member_stride;
            float fc[Q];
#pragma unroll
            for (int q = 0; q < Q; ++q) fc[q] = fk[q * p.nodes + node];
            float r, u, v;
            moments(fc, p.c, r, u, v);
            su += u; sv += v; sr += r;
            suu += double(u) * u; svv += double(v) * v; srr += double(r) * r;
        }
        const double iM = 1.0 / M;
        const double mu = su * iM, mv = sv * iM, mr = sr * iM;
        const double vu = fmax(suu * iM - mu * mu, 0.0), vv = fmax(svv * iM - mv * mv, 0.0),
                     vr = fmax(srr * iM - mr * mr, 0.0);
        const double du = mu - tu[node], dv = mv - tv[node], dr = mr - trho[node];
        acc[0] = du * du + dv * dv;
        acc[1] = vu + vv;
        acc[2] = dr * dr;
        acc[3] = vr;
    }
    for (int a = 0; a < 4; ++a) red[a][threadIdx.x] = acc[a];
    __syncthreads();
    for (int s = STATS_NT / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int a = 0; a < 4; ++a) red[a][threadIdx.x] += red[a][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int a = 0; a < 4; ++a) partials[blockIdx.x * 4 + a] = red[a][0];
}

inline int node_blocks(long long nodes, int nt) { return int((nodes + nt - 1) / nt); }

}  // namespace

bool lbm_fused_supported(int nx, int ny) { return nx % TX == 0 && ny % TY == 0; }

hipError_t launch_force_field(Ctx& c, int step_offset) {
    const int nt = 256;
    k_force_field<<<node_blocks(c.nodes, nt), nt, 0, c.stream>>>(
        c.d_force, c.cfg.force_seed, float(c.cfg.force_rand_stdev), c.d_step, step_offset,
        c.force_xy, c.nx, c.ny);
    return hipGetLastError();
}

hipError_t launch_step_advance(Ctx& c, int n) {
    k_step_advance<<<1, 1, 0, c.stream>>>(c.d_step, n);
    return hipGetLastError();
}

bool lbm_rows_supported(int nx, int ny, int h) {
    return nx % 64 == 0 && nx <= 1024 && (h == 8 || h == 16 || h == 32) && ny % h == 0 && ny >= 2;
}

// one step in full-row strips of h rows
static hipError_t launch_rows1(Ctx& c, const float* fin, float* fout, int h, bool force) {
    if (!lbm_rows_supported(c.nx, c.ny, h)) return hipErrorInvalidValue;
    const int strips = c.ny / h;
    const unsigned nb = unsigned(strips * c.M);
    const size_t lds = sizeof(float) * 8 * size_t(c.nx);
#define LBM_ROWS(HH)                                                                                       \
    (force ? k_lbm_rows<true, HH><<<nb, c.nx, lds, c.stream>>>(fin, fout, c.force_xy, c.lp, strips)      \
           : k_lbm_rows<false, HH><<<nb, c.nx, lds, c.stream>>>(fin, fout, c.force_xy, c.lp, strips))
    switch (h) {
        case 8: LBM_ROWS(8); break;
        case 16: LBM_ROWS(16); break;
        default: LBM_ROWS(32); break;
    }
#undef LBM_ROWS
    return hipGetLastError();
}

// LDS of the two-step strips: 120 B per column (12 velocity rows + 3 x 6 population
// rows of 4 B), 120 KB at nx = 1024 (the MI355X limit per workgroup is 160 KB)
bool lbm_rows2_supported(int nx, int ny, int h) {
    return nx % 64 == 0 && nx <= 1024 && (h == 16 || h == 32) && ny % h == 0 && ny >= 2;
}

// LDS of the three-step strips: 144 B per column (12 velocity rows + 2 x 12
// population rows of 4 B), 72 KB at nx = 512
bool lbm_rows3_supported(int nx, int ny, int h) {
    // byte offsets within one member's populations are 32-bit signed (buffer addressing)
    return nx % 64 == 0 && nx <= 512 && (h == 16 || h == 32) && ny % h == 0 && ny >= 2 &&
           (long long)nx * ny * Q * 4 < (1LL << 31);
}

int lbm_launch_steps(const Ctx& c) {
    const bool force = c.lp.has_force || c.lp.fric_over_dt != 0.f;
    if (force) return 1;
    if (c.lbm_tiles_per_block <= -LBM_THREE_STEP) return 3;
    return c.lbm_tiles_per_block <= -LBM_TWO_STEP ? 2 : 1;
}

static hipError_t launch_rows2(Ctx& c, const float* fin, float* fout, int h) {
    if (!lbm_rows2_supported(c.nx, c.ny, h)) return hipErrorInvalidValue;
    const int strips = c.ny / h;
    const unsigned nb = unsigned(strips * c.M);
    const size_t lds = sizeof(float) * (12 + 3 * P1Q) * size_t(c.nx);
    const void* fn = h == 16 ? reinterpret_cast<const void*>(&k_lbm_rows2<16>)
                             : reinterpret_cast<const void*>(&k_lbm_rows2<32>);
    if (lds > 65536) {  // beyond the default dynamic LDS of a workgroup
        const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
        if (e != hipSuccess) return e;
    }
    if (h == 16)
        k_lbm_rows2<16><<<nb, c.nx, lds, c.stream>>>(fin, fout, c.lp, strips);
    else
        k_lbm_rows2<32><<<nb, c.nx, lds, c.stream>>>(fin, fout, c.lp, strips);
    return hipGetLastError();
}

static hipError_t launch_rows3(Ctx& c, const float* fin, float* fout, int h) {
    if (!lbm_rows3_supported(c.nx, c.ny, h)) return hipErrorInvalidValue;
    const int strips = c.ny / h;
    const unsigned nb = unsigned(strips * c.M);
    const size_t lds = sizeof(float) * 36 * size_t(c.nx);
    const bool n256 = c.nx == 256;
    const void* fn = h == 16 ? (n256 ? reinterpret_cast<const void*>(&k_lbm_rows3<16, 256>)
                                     : reinterpret_cast<const void*>(&k_lbm_rows3<16, 0>))
                             : (n256 ? reinterpret_cast<const void*>(&k_lbm_rows3<32, 256>)
                                     : reinterpret_cast<const void*>(&k_lbm_rows3<32, 0>));
    if (lds > 65536) {
        const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
        if (e != hipSuccess) return e;
    }
    if (h == 16 && n256)
        k_lbm_rows3<16, 256><<<nb, c.nx, lds, c.stream>>>(fin, fout, c.lp, strips);
    else if (h == 16)
        k_lbm_rows3<16, 0><<<nb, c.nx, lds, c.stream>>>(fin, fout, c.lp, strips);
    else if (n256)
        k_lbm_rows3<32, 256><<<nb, c.nx, lds, c.stream>>>(fin, fout, c.lp, strips);
    else
        k_lbm_rows3<32, 0><<<nb, c.nx, lds, c.stream>>>(fin, fout, c.lp, strips);
    return hipGetLastError();
}

hipError_t launch_lbm_fused(Ctx& c, const float* fin, float* fout, int steps) {
    const bool force = c.lp.has_force || c.lp.fric_over_dt != 0.f;
    if (steps < 1 || steps > lbm_launch_steps(c)) return hipErrorInvalidValue;
    if (c.lbm_tiles_per_block <= -LBM_TWO_STEP) {
        // full-row strips of h rows taking three (or two) steps per launch, or one,
        // for the last steps of a forecast or a forced flow
        const int h = -c.lbm_tiles_per_block - (c.lbm_tiles_per_block <= -LBM_THREE_STEP ? LBM_THREE_STEP : LBM_TWO_STEP);
        if (steps == 3) return launch_rows3(c, fin, fout, h);
        if (steps == 2) return launch_rows2(c, fin, fout, h);
        return launch_rows1(c, fin, fout, h, force);
    }
    if (c.lbm_tiles_per_block < 0) return launch_rows1(c, fin, fout, -c.lbm_tiles_per_block, force);
    const int tiles_x = c.nx / TX, tiles_y = c.ny / TY;
    const int tiles = tiles_x * tiles_y;
    const int total = tiles * c.M;
    const int per_block = c.lbm_tiles_per_block;
    if (per_block < 1 || total % per_block != 0) return hipErrorInvalidValue;
    const int nb = total / per_block;
    if (force)
        k_lbm_fused<true><<<nb, NT, 0, c.stream>>>(fin, fout, c.force_xy, c.lp, tiles_x, tiles, per_block);
    else
        k_lbm_fused<false><<<nb, NT, 0, c.stream>>>(fin, fout, c.force_xy, c.lp, tiles_x, tiles, per_block);
    return hipGetLastError();
}

hipError_t launch_lbm_plain_member(Ctx& c, const float* fin, float* fout, int member) {
    const int nt = 256;
    const int nb = node_blocks(c.nodes, nt);
    const long long off = (long long)member * c.lp.member_stride;
    k_stream_macro<<<nb, nt, 0, c.stream>>>(fin + off, fout + off, c.plain_ruv, c.lp);
    k_les<<<nb, nt, 0, c.stream>>>(c.plain_ruv, c.plain_nus, c.lp);
    k_collide<<<nb, nt, 0, c.stream>>>(fout + off, c.plain_ruv, c.plain_nus, c.force_xy, c.lp);
    return hipGetLastError();
}

hipError_t launch_set_feq(Ctx& c, int member, const float* d_rho, const float* d_u,
                          const float* d_v) {
    const int nt = 256;
    k_set_feq<<<node_blocks(c.nodes, nt), nt, 0, c.stream>>>(
        c.f[c.cur] + (long long)member * c.lp.member_stride, d_rho, d_u, d_v, c.lp);
    return hipGetLastError();
}

hipError_t launch_nudge(Ctx& c, const float* ro, const float* uo, const float* vo, float alpha, int xyprune,
                        int skip_east, int skip_north, int use_obs_rho) {
    const int nt = 256;
    k_nudge<<<node_blocks(c.nodes, nt), nt, 0, c.stream>>>(c.f[c.cur], ro, uo, vo, c.lp, alpha, xyprune, skip_east,
                                                          skip_north, use_obs_rho);
    return hipGetLastError();
}

hipError_t launch_perturb(Ctx& c, int member, const float* ur, const float* vr, float eps) {
    const int nt = 256;
    k_perturb<<<node_blocks(c.nodes, nt), nt, 0, c.stream>>>(c.f[c.cur] + (long long)member * c.lp.member_stride,
                                                            ur, vr, c.lp, eps);
    return hipGetLastError();
}

hipError_t launch_macro(Ctx& c, const float* f_member, float* rho, float* u, float* v) {
    const int nt = 256;
    k_macro<<<node_blocks(c.nodes, nt), nt, 0, c.stream>>>(f_member, rho, u, v, c.lp);
    return hipGetLastError();
}

int ensemble_stats_blocks(const Ctx& c) { return node_blocks(c.nodes, STATS_NT); }

// The partial sums of lbmletkf_ensemble_stats_noalloc live in this file-scope array
// instead of a buffer the library allocates: 4 doubles per block of k_ensemble_stats,
// for grids of up to STATS_SCRATCH_BLOCKS * STATS_NT = 2^24 nodes (2 MiB per device,
// held until the process ends).  Every call on a device shares that device's copy
// and overwrites the slots it uses, so calls that overlap on one device race.
static __device__ __align__(256) double scratch[STATS_SCRATCH_BLOCKS * 4];

hipError_t stats_scratch_address(double** out) {
    // the device address of `scratch` on the current device, looked up once per device
    // (the host-side name is not a device address)
    static std::mutex mu;
    static std::map<int, double*> addr;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lock(mu);
    auto it = addr.find(dev);
    if (it == addr.end()) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        it = addr.emplace(dev, static_cast<double*>(p)).first;
    }
    *out = it->second;
    return hipSuccess;
}

hipError_t launch_ensemble_fields(Ctx& c, float* uv, float* out) {
    const int nt = 256;
    const dim3 grid_uv(node_blocks(c.nodes, nt), unsigned(c.M));
    k_member_uv<<<grid_uv, nt, 0, c.stream>>>(c.f[c.cur], uv, c.lp);
    k_ensemble_fields<<<node_blocks(c.nodes, nt), nt, 0, c.stream>>>(uv, out, c.M, float(c.cfg.dx), c.lp);
    return hipGetLastError();
}

hipError_t launch_ensemble_stats(Ctx& c, const float* truth_rho, const float* truth_u,
                                 const float* truth_v, double* partials, int n_blocks) {
    k_ensemble_stats<<<n_blocks, STATS_NT, 0, c.stream>>>(c.f[c.cur], truth_rho, truth_u,
                                                          truth_v, partials, c.lp);
    return hipGetLastError();
}

}  // namespace lbmletkf
