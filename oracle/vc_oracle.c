/*
 * vc_oracle.c -- CPU restatement of the Voxel Cores volumetric hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the timed CPU baseline).  The
 * product library never links or loads this file.  Deliberately simple,
 * single-threaded code: every function is written for obviousness, not speed.
 *
 * Layout: uint8 occupancy, C order [x][y][z] (z fastest), i = (x*ny+y)*nz+z.
 */
#include "vc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX(x, y, z) (((int64_t)(x) * ny + (y)) * nz + (z))

static int64_t floordiv64(int64_t a, int64_t b) /* b > 0 */
{
    int64_t q = a / b;
    if ((a % b) != 0 && a < 0)
        q--;
    return q;
}

/* ======================================================================
 * Exact squared EDT + feature transform (project-defined; no reference
 * counterpart).  Meijster-style lower envelope of parabolas per axis, ties
 * resolved towards the lower index; axis order z, y, x so that the final
 * minimiser is the lowest linear index (x most significant).
 * ==================================================================== */
void or_envelope_1d(const int64_t* f, int n, int64_t* D, int32_t* arg)
{
    int32_t* s = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int64_t* t = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    int q = -1;
    for (int u = 0; u < n; ++u)
    {
        if (f[u] < 0)
            continue;
        while (q >= 0)
        {
            int64_t tq = t[q];
            int64_t i = s[q];
            int64_t Fi = (tq - i) * (tq - i) + f[i];
            int64_t Fu = (tq - u) * (tq - u) + f[u];
            if (Fi > Fu)
                q--;
            else
                break;
        }
        if (q < 0)
        {
            q = 0;
            s[0] = u;
            t[0] = 0;
        }
        else
        {
            int64_t i = s[q];
            int64_t num = (int64_t)u * u - i * i + f[u] - f[i];
            int64_t w = 1 + floordiv64(num, 2 * ((int64_t)u - i));
            if (w < n)
            {
                q++;
                s[q] = u;
                t[q] = w;
            }
        }
    }
    if (q < 0)
    {
        for (int u = 0; u < n; ++u)
        {
            D[u] = -1;
            arg[u] = -1;
        }
    }
    else
    {
        for (int u = n - 1; u >= 0; --u)
        {
            int64_t i = s[q];
            D[u] = ((int64_t)u - i) * ((int64_t)u - i) + f[i];
            arg[u] = (int32_t)i;
            if (u == t[q])
                q--;
        }
    }
    free(s);
    free(t);
}

#ifdef _OPENMP
#include <omp.h>
#endif

/* threads of the OpenMP loops (1 = the sequential restatement); returns
 * the number in use */
int or_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0)
        omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int or_edt_ft(const uint8_t* occ, int nx, int ny, int nz, int32_t* d2,
              int32_t* ft)
{
    if (nx <= 0 || ny <= 0 || nz <= 0)
        return -1;
    int64_t N = (int64_t)nx * ny * nz;
    int nmax = nx > ny ? nx : ny;
    nmax = nmax > nz ? nmax : nz;
    int32_t* zq = (int32_t*)malloc(sizeof(int32_t) * N);
    int32_t* yq = (int32_t*)malloc(sizeof(int32_t) * N);
    int32_t* zz = (int32_t*)malloc(sizeof(int32_t) * N);
    /* rows and columns are independent: with OpenMP each thread takes whole
     * rows with its own scratch */
#pragma omp parallel
    {
    int64_t* f = (int64_t*)malloc(sizeof(int64_t) * nmax);
    int64_t* D = (int64_t*)malloc(sizeof(int64_t) * nmax);
    int32_t* a = (int32_t*)malloc(sizeof(int32_t) * nmax);
    /* pass z: nearest background along the row */
#pragma omp for schedule(static)
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
        {
            for (int z = 0; z < nz; ++z)
                f[z] = occ[IDX(x, y, z)] ? -1 : 0;
            or_envelope_1d(f, nz, D, a);
            for (int z = 0; z < nz; ++z)
                zq[IDX(x, y, z)] = a[z];
        }
    /* pass y */
#pragma omp for schedule(static)
    for (int x = 0; x < nx; ++x)
        for (int z = 0; z < nz; ++z)
        {
            for (int y = 0; y < ny; ++y)
            {
                int32_t q = zq[IDX(x, y, z)];
                f[y] = q < 0 ? -1 : (int64_t)(z - q) * (z - q);
            }
            or_envelope_1d(f, ny, D, a);
            for (int y = 0; y < ny; ++y)
            {
                yq[IDX(x, y, z)] = a[y];
                zz[IDX(x, y, z)] = a[y] < 0 ? -1 : zq[IDX(x, a[y], z)];
            }
        }
    /* pass x */
#pragma omp for schedule(static)
    for (int y = 0; y < ny; ++y)
        for (int z = 0; z < nz; ++z)
        {
            for (int x = 0; x < nx; ++x)
            {
                int64_t i = IDX(x, y, z);
                f[x] = yq[i] < 0 ? -1
                                 : (int64_t)(y - yq[i]) * (y - yq[i]) +
                                       (int64_t)(z - zz[i]) * (z - zz[i]);
            }
            or_envelope_1d(f, nx, D, a);
            for (int x = 0; x < nx; ++x)
            {
                int64_t i = IDX(x, y, z);
                if (a[x] < 0)
                {
                    d2[i] = -1;
                    ft[i] = -1;
                }
                else
                {
                    int64_t j = IDX(a[x], y, z);
                    d2[i] = (int32_t)D[x];
                    ft[i] = (int32_t)IDX(a[x], yq[j], zz[j]);
                }
            }
        }
    free(f);
    free(D);
    free(a);
    }
    free(zq);
    free(yq);
    free(zz);
    return 0;
}

int or_edt_brute(const uint8_t* occ, int nx, int ny, int nz, int32_t* d2,
                 int32_t* ft)
{
    int64_t N = (int64_t)nx * ny * nz;
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z)
            {
                int64_t best = -1, barg = -1;
                for (int64_t q = 0; q < N; ++q)
                {
                    if (occ[q])
                        continue;
                    int qz = (int)(q % nz);
                    int qy = (int)((q / nz) % ny);
                    int qx = (int)(q / ((int64_t)nz * ny));
                    int64_t d = (int64_t)(x - qx) * (x - qx) +
                                (int64_t)(y - qy) * (y - qy) +
                                (int64_t)(z - qz) * (z - qz);
                    if (best < 0 || d < best)
                    {
                        best = d;
                        barg = q;
                    }
                }
                d2[IDX(x, y, z)] = (int32_t)best;
                ft[IDX(x, y, z)] = (int32_t)barg;
            }
    return 0;
}

/* ======================================================================
 * Per-voxel medial measure (project-defined R16)
 * ==================================================================== */
void or_lambda2(const uint8_t* occ, const int32_t* ft, int nx, int ny, int nz,
                int32_t* lam2)
{
    static const int O6[6][3] = {{-1, 0, 0}, {1, 0, 0},  {0, -1, 0},
                                 {0, 1, 0},  {0, 0, -1}, {0, 0, 1}};
#pragma omp parallel for schedule(static)
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z)
            {
                int64_t i = IDX(x, y, z);
                int32_t m = 0;
                if (occ[i] && ft[i] >= 0)
                {
                    int64_t f = ft[i];
                    int fz = (int)(f % nz), fy = (int)((f / nz) % ny),
                        fx = (int)(f / ((int64_t)nz * ny));
                    for (int o = 0; o < 6; ++o)
                    {
                        int a = x + O6[o][0], b = y + O6[o][1],
                            c = z + O6[o][2];
                        if (a < 0 || b < 0 || c < 0 || a >= nx || b >= ny ||
                            c >= nz)
                            continue;
                        int64_t j = IDX(a, b, c);
                        if (!occ[j] || ft[j] < 0)
                            continue;
                        int64_t g = ft[j];
                        int gz = (int)(g % nz), gy = (int)((g / nz) % ny),
                            gx = (int)(g / ((int64_t)nz * ny));
                        int32_t d = (fx - gx) * (fx - gx) +
                                    (fy - gy) * (fy - gy) +
                                    (fz - gz) * (fz - gz);
                        if (d > m)
                            m = d;
                    }
                }
                lam2[i] = m;
            }
}

int64_t or_level_threshold(float level)
{
    double l = (double)level;
    if (!(l > 0.0))
        return 0;
    double l2 = l * l; /* exact: 24-bit mantissa squared fits in 53 bits */
    if (l2 > 9.0e18)
        return INT64_MAX;
    return (int64_t)ceil(l2);
}

void or_levels(const uint8_t* occ, const int32_t* lam2, int64_t n,
               const int64_t* T, int nlev, uint8_t* lvl)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
    {
        uint8_t v = 255;
        if (occ[i])
            for (int k = 0; k < nlev; ++k)
                if ((int64_t)lam2[i] < T[k])
                {
                    v = (uint8_t)k;
                    break;
                }
        lvl[i] = v;
    }
}

/* ======================================================================
 * Simple points for (26,6) adjacency.
 * Cube cell j = (dz+1) + 3(dy+1) + 9(dx+1); centre j = 13.
 * ==================================================================== */
static void cell_xyz(int j, int* dx, int* dy, int* dz)
{
    *dz = j % 3 - 1;
    *dy = (j / 3) % 3 - 1;
    *dx = j / 9 - 1;
}

static uint32_t nb26_to_cube(uint32_t nb26)
{
    uint32_t lo = nb26 & 0x1FFFu;          /* cells 0..12 */
    uint32_t hi = (nb26 >> 13) & 0x1FFFu;  /* cells 14..26 */
    return lo | (hi << 14);
}

/* Adjacency of the 27 cells as one 27-bit mask per cell, from coordinate
 * differences (26: Chebyshev distance 1; 6: L1 distance 1).  Built before
 * any parallel loop that tests simplicity. */
static uint32_t g_adj26[27], g_adj6[27];
static int g_adj_ready = 0;

static void init_adj(void)
{
    if (g_adj_ready)
        return;
    for (int u = 0; u < 27; ++u)
    {
        int ux, uy, uz;
        cell_xyz(u, &ux, &uy, &uz);
        uint32_t a26 = 0, a6 = 0;
        for (int v = 0; v < 27; ++v)
        {
            int vx, vy, vz;
            cell_xyz(v, &vx, &vy, &vz);
            int ax = abs(ux - vx), ay = abs(uy - vy), az = abs(uz - vz);
            if (v != u && ax <= 1 && ay <= 1 && az <= 1)
                a26 |= 1u << v;
            if (ax + ay + az == 1)
                a6 |= 1u << v;
        }
        g_adj26[u] = a26;
        g_adj6[u] = a6;
    }
    g_adj_ready = 1;
}

/* BFS component count of `set` (27-bit) under adjacency `adj` (6 or 26),
 * one cell popped at a time; if `must` != 0 only components meeting `must`
 * are counted. */
static int bfs_components(uint32_t set, int adj, uint32_t must)
{
    const uint32_t* nbr = adj == 26 ? g_adj26 : g_adj6;
    int count = 0;
    uint32_t rest = set;
    while (rest)
    {
        uint32_t comp = rest & (~rest + 1u);
        uint32_t queue = comp;
        while (queue)
        {
            int u = __builtin_ctz(queue);
            queue &= queue - 1u;
            uint32_t fresh = nbr[u] & set & ~comp;
            comp |= fresh;
            queue |= fresh;
        }
        rest &= ~comp;
        if (!must || (comp & must))
            count++;
    }
    return count;
}

int or_is_simple(uint32_t nb26)
{
    init_adj();
    uint32_t X = nb26_to_cube(nb26); /* centre bit 13 clear */
    uint32_t n18 = 0, n6 = 0;
    for (int j = 0; j < 27; ++j)
    {
        int dx, dy, dz;
        cell_xyz(j, &dx, &dy, &dz);
        int l1 = abs(dx) + abs(dy) + abs(dz);
        if (l1 >= 1 && l1 <= 2)
            n18 |= 1u << j;
        if (l1 == 1)
            n6 |= 1u << j;
    }
    int t26 = bfs_components(X, 26, 0);
    if (t26 != 1)
        return 0;
    uint32_t B = (~X) & n18;
    int t6 = bfs_components(B, 6, n6);
    return t6 == 1;
}

/* ---- topological definition on a 5x5x5 box (independent check) -------- */
static int64_t box_euler(const uint8_t* b, int n)
{
    int m = 2 * n + 1;
    uint8_t* mark = (uint8_t*)calloc((size_t)m * m * m, 1);
    for (int x = 0; x < n; ++x)
        for (int y = 0; y < n; ++y)
            for (int z = 0; z < n; ++z)
                if (b[(x * n + y) * n + z])
                    for (int a = 0; a < 3; ++a)
                        for (int c = 0; c < 3; ++c)
                            for (int d = 0; d < 3; ++d)
                                mark[((2 * x + a) * m + 2 * y + c) * m + 2 * z +
                                     d] = 1;
    int64_t chi = 0;
    for (int x = 0; x < m; ++x)
        for (int y = 0; y < m; ++y)
            for (int z = 0; z < m; ++z)
                if (mark[(x * m + y) * m + z])
                {
                    int dim = (x & 1) + (y & 1) + (z & 1);
                    chi += (dim & 1) ? -1 : 1;
                }
    free(mark);
    return chi;
}

static int box_components(const uint8_t* b, int n, int want, int adj)
{
    int N = n * n * n;
    int* seen = (int*)calloc(N, sizeof(int));
    int* q = (int*)malloc(sizeof(int) * N);
    int count = 0;
    for (int s = 0; s < N; ++s)
    {
        if ((b[s] != 0) != want || seen[s])
            continue;
        count++;
        int h = 0, t = 0;
        q[t++] = s;
        seen[s] = 1;
        while (h < t)
        {
            int u = q[h++];
            int ux = u / (n * n), uy = (u / n) % n, uz = u % n;
            for (int dx = -1; dx <= 1; ++dx)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dz = -1; dz <= 1; ++dz)
                    {
                        int l1 = abs(dx) + abs(dy) + abs(dz);
                        if (l1 == 0 || (adj == 6 && l1 != 1))
                            continue;
                        int vx = ux + dx, vy = uy + dy, vz = uz + dz;
                        if (vx < 0 || vy < 0 || vz < 0 || vx >= n ||
                            vy >= n || vz >= n)
                            continue;
                        int v = (vx * n + vy) * n + vz;
                        if ((b[v] != 0) != want || seen[v])
                            continue;
                        seen[v] = 1;
                        q[t++] = v;
                    }
        }
    }
    free(seen);
    free(q);
    return count;
}

int or_is_simple_topo(uint32_t nb26)
{
    uint8_t box[125];
    memset(box, 0, sizeof(box));
    uint32_t X = nb26_to_cube(nb26) | (1u << 13);
    for (int j = 0; j < 27; ++j)
        if ((X >> j) & 1u)
        {
            int dx, dy, dz;
            cell_xyz(j, &dx, &dy, &dz);
            box[((dx + 2) * 5 + dy + 2) * 5 + dz + 2] = 1;
        }
    int64_t chi0 = box_euler(box, 5);
    int c0 = box_components(box, 5, 1, 26);
    int b0 = box_components(box, 5, 0, 6);
    box[(2 * 5 + 2) * 5 + 2] = 0;
    int64_t chi1 = box_euler(box, 5);
    int c1 = box_components(box, 5, 1, 26);
    int b1 = box_components(box, 5, 0, 6);
    return chi0 == chi1 && c0 == c1 && b0 == b1;
}

/* ---- bit-parallel (26,6) test used to fill the full table ------------ */
#define CUBE_ALL 0x7FFFFFFu
static uint32_t g_zlo, g_zhi, g_ylo, g_yhi; /* cells with z=0/2, y=0/2 */
static uint32_t g_n6, g_n18, g_n26;
static int g_masks_ready = 0;

static void init_masks(void)
{
    if (g_masks_ready)
        return;
    uint32_t zlo = 0, zhi = 0, ylo = 0, yhi = 0, n6 = 0, n18 = 0, n26 = 0;
    for (int j = 0; j < 27; ++j)
    {
        int dx, dy, dz;
        cell_xyz(j, &dx, &dy, &dz);
        if (dz == -1)
            zlo |= 1u << j;
        if (dz == 1)
            zhi |= 1u << j;
        if (dy == -1)
            ylo |= 1u << j;
        if (dy == 1)
            yhi |= 1u << j;
        int l1 = abs(dx) + abs(dy) + abs(dz);
        if (l1 == 1)
            n6 |= 1u << j;
        if (l1 >= 1 && l1 <= 2)
            n18 |= 1u << j;
        if (l1 >= 1)
            n26 |= 1u << j;
    }
    g_zlo = zlo;
    g_zhi = zhi;
    g_ylo = ylo;
    g_yhi = yhi;
    g_n6 = n6;
    g_n18 = n18;
    g_n26 = n26;
    g_masks_ready = 1;
}

static uint32_t dil6(uint32_t m)
{
    uint32_t r = m;
    r |= (m << 1) & ~g_zlo;
    r |= (m >> 1) & ~g_zhi;
    r |= (m << 3) & ~g_ylo;
    r |= (m >> 3) & ~g_yhi;
    r |= (m << 9);
    r |= (m >> 9);
    return r & CUBE_ALL;
}

static uint32_t dil26(uint32_t m)
{
    uint32_t r = (m | ((m << 1) & ~g_zlo) | ((m >> 1) & ~g_zhi)) & CUBE_ALL;
    r = (r | ((r << 3) & ~g_ylo) | ((r >> 3) & ~g_yhi)) & CUBE_ALL;
    r = r | (r << 9) | (r >> 9);
    return r & CUBE_ALL;
}

static int fast_simple(uint32_t nb26)
{
    uint32_t X = nb26_to_cube(nb26) & g_n26;
    if (!X)
        return 0;
    /* T26 == 1 */
    uint32_t seed = X & (~X + 1u);
    uint32_t c = seed, prev;
    do
    {
        prev = c;
        c = dil26(c) & X;
    } while (c != prev);
    if (c != X)
        return 0;
    /* T6 == 1 over background in N18 */
    uint32_t B = (~X) & g_n18;
    uint32_t f6 = B & g_n6;
    if (!f6)
        return 0;
    seed = f6 & (~f6 + 1u);
    c = seed;
    do
    {
        prev = c;
        c = dil6(c) & B;
    } while (c != prev);
    return (f6 & ~c) == 0;
}

void or_simple_lut(uint32_t* lut_words)
{
    init_masks();
    for (uint32_t w = 0; w < (1u << 21); ++w)
    {
        uint32_t word = 0;
        for (uint32_t b = 0; b < 32; ++b)
            if (fast_simple((w << 5) | b))
                word |= 1u << b;
        lut_words[w] = word;
    }
}

/* All 2^26 neighbourhood codes: the BFS definition (or_is_simple) against
 * the bit-parallel one that fills the table (fast_simple).  Returns the
 * number of codes where they differ; *n_simple gets the number of simple
 * codes by the BFS definition.  With `topo_stride` > 0, every
 * topo_stride-th code (starting at topo_offset) is also checked against the
 * topological definition on a 5x5x5 box (or_is_simple_topo) and the
 * disagreements are added to *topo_mismatch. */
int64_t or_simple_exhaustive(int64_t* n_simple, int64_t topo_stride,
                             int64_t topo_offset, int64_t* topo_mismatch,
                             int64_t* topo_checked)
{
    init_masks();
    init_adj();
    int64_t bad = 0, simple = 0, tbad = 0, tchk = 0;
#pragma omp parallel for schedule(dynamic, 65536) reduction(+ : bad, simple, tbad, tchk)
    for (int64_t c = 0; c < ((int64_t)1 << 26); ++c)
    {
        int a = or_is_simple((uint32_t)c);
        int b = fast_simple((uint32_t)c);
        simple += a;
        bad += a != b;
        if (topo_stride > 0 && c % topo_stride == topo_offset % topo_stride)
        {
            tchk++;
            tbad += a != or_is_simple_topo((uint32_t)c);
        }
    }
    if (n_simple)
        *n_simple = simple;
    if (topo_mismatch)
        *topo_mismatch = tbad;
    if (topo_checked)
        *topo_checked = tchk;
    return bad;
}

/* ======================================================================
 * Measure-ordered subfield thinning (project-defined R15).
 * For each level k (ascending): repeat full sweeps until a sweep removes
 * nothing; a sweep visits the 8 subfields s = (x&1)|(y&1)<<1|(z&1)<<2 in
 * ascending order and removes every alive voxel of the subfield whose level
 * index is <= k and which is simple in the current image.
 * ==================================================================== */
static uint8_t* g_memo = NULL; /* per config: 0 unknown, 1 no, 2 yes */

/* threads that meet the same configuration store the same value */
static int memo_simple(uint32_t nb)
{
    uint8_t v = __atomic_load_n(&g_memo[nb], __ATOMIC_RELAXED);
    if (!v)
    {
        v = or_is_simple(nb) ? 2 : 1;
        __atomic_store_n(&g_memo[nb], v, __ATOMIC_RELAXED);
    }
    return v == 2;
}

static uint32_t gather26(const uint8_t* label, int nx, int ny, int nz, int x,
                         int y, int z)
{
    uint32_t nb = 0;
    int b = 0;
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz)
            {
                if (dx == 0 && dy == 0 && dz == 0)
                    continue;
                int a = x + dx, c = y + dy, d = z + dz;
                int in = a >= 0 && c >= 0 && d >= 0 && a < nx && c < ny &&
                         d < nz && label[IDX(a, c, d)] == 1;
                if (in)
                    nb |= 1u << b;
                b++;
            }
    return nb;
}

int or_thin(uint8_t* label, const uint8_t* lvl, int nx, int ny, int nz,
            int nlev, or_thin_stats* st)
{
    if (nlev < 0 || nlev > 64)
        return -1;
    if (st)
        memset(st, 0, sizeof(*st));
    if (!g_memo)
        g_memo = (uint8_t*)calloc((size_t)1 << 26, 1);
    init_adj();
    for (int k = 0; k < nlev; ++k)
    {
        int changed;
        do
        {
            changed = 0;
            if (st)
                st->sweeps[k]++;
            for (int s = 0; s < 8; ++s)
            {
                int sx = s & 1, sy = (s >> 1) & 1, sz = (s >> 2) & 1;
                int64_t tests = 0, removed = 0;
                /* voxels of one subfield are never 26-adjacent: a pass reads
                 * no label it writes, so its voxels are independent */
#pragma omp parallel for schedule(dynamic, 1) reduction(| : changed) \
    reduction(+ : tests, removed)
                for (int x = sx; x < nx; x += 2)
                    for (int y = sy; y < ny; y += 2)
                        for (int z = sz; z < nz; z += 2)
                        {
                            int64_t i = IDX(x, y, z);
                            if (label[i] != 1 || lvl[i] > k)
                                continue;
                            uint32_t nb = gather26(label, nx, ny, nz, x, y, z);
                            tests++;
                            if (memo_simple(nb))
                            {
                                label[i] = (uint8_t)(2 + k);
                                changed = 1;
                                removed++;
                            }
                        }
                if (st)
                {
                    st->tests += tests;
                    st->removed[k] += removed;
                }
            }
        } while (changed);
    }
    return 0;
}

/* ======================================================================
 * Topology of voxel sets
 * ==================================================================== */
static int occ_at(const uint8_t* occ, int nx, int ny, int nz, int x, int y,
                  int z)
{
    if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz)
        return 0;
    return occ[IDX(x, y, z)] != 0;
}

int64_t or_euler_voxels(const uint8_t* occ, int nx, int ny, int nz)
{
    int64_t V = 0, E = 0, F = 0, C = 0;
    /* vertices: lattice points with any of the 8 incident voxels inside */
#pragma omp parallel for schedule(static) reduction(+ : V)
    for (int x = 0; x <= nx; ++x)
        for (int y = 0; y <= ny; ++y)
            for (int z = 0; z <= nz; ++z)
            {
                int any = 0;
                for (int a = -1; a <= 0 && !any; ++a)
                    for (int b = -1; b <= 0 && !any; ++b)
                        for (int c = -1; c <= 0 && !any; ++c)
                            any |= occ_at(occ, nx, ny, nz, x + a, y + b, z + c);
                V += any;
            }
    /* edges along each axis: 4 incident voxels */
#pragma omp parallel for schedule(static) reduction(+ : E)
    for (int x = 0; x < nx; ++x) /* x-edges at (x, y, z) lattice, y<=ny, z<=nz */
        for (int y = 0; y <= ny; ++y)
            for (int z = 0; z <= nz; ++z)
            {
                int any = 0;
                for (int b = -1; b <= 0; ++b)
                    for (int c = -1; c <= 0; ++c)
                        any |= occ_at(occ, nx, ny, nz, x, y + b, z + c);
                E += any;
            }
#pragma omp parallel for schedule(static) reduction(+ : E)
    for (int x = 0; x <= nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z <= nz; ++z)
            {
                int any = 0;
                for (int a = -1; a <= 0; ++a)
                    for (int c = -1; c <= 0; ++c)
                        any |= occ_at(occ, nx, ny, nz, x + a, y, z + c);
                E += any;
            }
#pragma omp parallel for schedule(static) reduction(+ : E)
    for (int x = 0; x <= nx; ++x)
        for (int y = 0; y <= ny; ++y)
            for (int z = 0; z < nz; ++z)
            {
                int any = 0;
                for (int a = -1; a <= 0; ++a)
                    for (int b = -1; b <= 0; ++b)
                        any |= occ_at(occ, nx, ny, nz, x + a, y + b, z);
                E += any;
            }
    /* faces: 2 incident voxels */
#pragma omp parallel for schedule(static) reduction(+ : F)
    for (int x = 0; x <= nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z)
                F += occ_at(occ, nx, ny, nz, x - 1, y, z) |
                     occ_at(occ, nx, ny, nz, x, y, z);
#pragma omp parallel for schedule(static) reduction(+ : F)
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y <= ny; ++y)
            for (int z = 0; z < nz; ++z)
                F += occ_at(occ, nx, ny, nz, x, y - 1, z) |
                     occ_at(occ, nx, ny, nz, x, y, z);
#pragma omp parallel for schedule(static) reduction(+ : F)
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z <= nz; ++z)
                F += occ_at(occ, nx, ny, nz, x, y, z - 1) |
                     occ_at(occ, nx, ny, nz, x, y, z);
#pragma omp parallel for schedule(static) reduction(+ : C)
    for (int64_t i = 0; i < (int64_t)nx * ny * nz; ++i)
        C += occ[i] != 0;
    return V - E + F - C;
}

/* Connected components of the voxels whose occupancy equals `want`, by BFS
 * over a copy of the grid padded by one voxel on every side: the pad holds
 * `pad` (1 = member, 0 = not), so no neighbour needs a bounds check.  With
 * pad = 1 every component touching the grid border is joined through the
 * pad into one component (and the pad alone counts as one when none does). */
static int64_t padded_components(const uint8_t* occ, int nx, int ny, int nz,
                                 int want, int adj, int pad)
{
    const int64_t px = nx + 2, py = ny + 2, pz = nz + 2;
    const int64_t NP = px * py * pz;
    uint8_t* m = (uint8_t*)malloc((size_t)NP);
    int64_t* q = (int64_t*)malloc(sizeof(int64_t) * (size_t)NP);
    if (!m || !q)
    {
        free(m);
        free(q);
        return -1;
    }
    memset(m, pad ? 1 : 0, (size_t)NP);
#pragma omp parallel for schedule(static)
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
        {
            uint8_t* row = m + ((x + 1) * py + (y + 1)) * pz + 1;
            const uint8_t* src = occ + IDX(x, y, 0);
            for (int z = 0; z < nz; ++z)
                row[z] = (uint8_t)((src[z] != 0) == want);
        }
    int64_t off[26];
    int noff = 0;
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz)
            {
                int l1 = abs(dx) + abs(dy) + abs(dz);
                if (l1 == 0 || (adj == 6 && l1 != 1))
                    continue;
                off[noff++] = (dx * py + dy) * pz + dz;
            }
    /* members are 1; a visited member becomes 2.  Neighbours of a
     * non-pad cell stay inside the padded grid; pad cells (pad = 1) have
     * out-of-range neighbours, so their offsets are checked against the
     * padded bounds. */
    int64_t count = 0;
    for (int64_t s = 0; s < NP; ++s)
    {
        if (m[s] != 1)
            continue;
        count++;
        int64_t h = 0, t = 0;
        q[t++] = s;
        m[s] = 2;
        while (h < t)
        {
            int64_t u = q[h++];
            int64_t uz = u % pz, uy = (u / pz) % py, ux = u / (pz * py);
            int edge = ux == 0 || uy == 0 || uz == 0 || ux == px - 1 ||
                       uy == py - 1 || uz == pz - 1;
            if (!edge)
            {
                for (int k = 0; k < noff; ++k)
                {
                    int64_t v = u + off[k];
                    if (m[v] == 1)
                    {
                        m[v] = 2;
                        q[t++] = v;
                    }
                }
                continue;
            }
            for (int dx = -1; dx <= 1; ++dx)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dz = -1; dz <= 1; ++dz)
                    {
                        int l1 = abs(dx) + abs(dy) + abs(dz);
                        if (l1 == 0 || (adj == 6 && l1 != 1))
                            continue;
                        int64_t vx = ux + dx, vy = uy + dy, vz = uz + dz;
                        if (vx < 0 || vy < 0 || vz < 0 || vx >= px ||
                            vy >= py || vz >= pz)
                            continue;
                        int64_t v = (vx * py + vy) * pz + vz;
                        if (m[v] == 1)
                        {
                            m[v] = 2;
                            q[t++] = v;
                        }
                    }
        }
    }
    free(m);
    free(q);
    return count;
}

int64_t or_components26(const uint8_t* occ, int nx, int ny, int nz)
{
    return padded_components(occ, nx, ny, nz, 1, 26, 0);
}

int64_t or_bg_components6(const uint8_t* occ, int nx, int ny, int nz)
{
    /* the outside of the grid is background: components touching the grid
     * border are one (the unbounded one) */
    return padded_components(occ, nx, ny, nz, 0, 6, 1);
}

/* ======================================================================
 * Surfacer restatement (reference-pinned): src/surfacing.cpp:49-660,
 * include/surfacing.h:97-240, include/spaceinfo.h:122-174.
 * ==================================================================== */
/* surface components of a 2x2x2 mask: corner i <-> mask bit (7-i), corners
 * adjacent iff their indices differ in one bit.  Equals the number of
 * connected edge-sets of the marching-cubes cycle table for that mask
 * (src/surfacing.cpp:466-540): (6-components of inside corners) +
 * (18-components of outside corners) - 1, with masks 0/255 and the four
 * antipodal-hole masks (src/surfacing.cpp:545-550) set to 0. */
static int corner_components(int set, int adj18)
{
    int seen = 0, n = 0;
    for (int s = 0; s < 8; ++s)
    {
        if (!((set >> s) & 1) || ((seen >> s) & 1))
            continue;
        n++;
        int stack[8], sp = 0;
        stack[sp++] = s;
        seen |= 1 << s;
        while (sp)
        {
            int u = stack[--sp];
            for (int v = 0; v < 8; ++v)
            {
                if (!((set >> v) & 1) || ((seen >> v) & 1))
                    continue;
                int d = u ^ v, pc = __builtin_popcount(d);
                if (pc == 1 || (adj18 && pc == 2))
                {
                    seen |= 1 << v;
                    stack[sp++] = v;
                }
            }
        }
    }
    return n;
}

void or_cube_components(int32_t* tbl)
{
    for (int mask = 0; mask < 256; ++mask)
    {
        int in = 0, out = 0;
        for (int i = 0; i < 8; ++i)
        {
            if ((mask >> (7 - i)) & 1)
                in |= 1 << i;
            else
                out |= 1 << i;
        }
        if (mask == 0 || mask == 255)
        {
            tbl[mask] = 0;
            continue;
        }
        tbl[mask] = corner_components(in, 0) + corner_components(out, 1) - 1;
    }
    tbl[126] = tbl[189] = tbl[219] = tbl[231] = 0;
}

static const int S_OFF[6][3] = {{-1, 0, 0}, {1, 0, 0},  {0, -1, 0},
                                {0, 1, 0},  {0, 0, -1}, {0, 0, 1}};
static const int S_CWRT[6][4] = {{2, 3, 7, 6}, {0, 1, 5, 4}, {4, 6, 2, 0},
                                 {1, 3, 7, 5}, {0, 2, 3, 1}, {5, 7, 6, 4}};
static const int S_FWRT[6] = {3, 2, 4, 5, 0, 1};
static const int S_FCRN[6][4] = {{0, 1, 3, 2}, {4, 5, 7, 6}, {0, 4, 5, 1},
                                 {2, 6, 7, 3}, {4, 0, 2, 6}, {5, 1, 3, 7}};
/* corner c of voxel (i,j,k) sits at (i,j,k) + 0.5*S_CSGN[c] */
static const int S_CSGN[8][3] = {{1, -1, -1}, {1, 1, -1},  {-1, -1, -1},
                                 {-1, 1, -1}, {1, -1, 1},  {1, 1, 1},
                                 {-1, -1, 1}, {-1, 1, 1}};

int or_boundary(const uint8_t* occ, int nx, int ny, int nz, float* sites,
                int64_t site_cap, int64_t* n_sites, int32_t* quads,
                int64_t quad_cap, int64_t* n_quads, or_euler* eu)
{
    int64_t N = (int64_t)nx * ny * nz;
    int64_t cx1 = nx + 1, cy1 = ny + 1, cz1 = nz + 1;
    int64_t NC = cx1 * cy1 * cz1;
    int32_t* cmap = (int32_t*)malloc(sizeof(int32_t) * NC);
    for (int64_t i = 0; i < NC; ++i)
        cmap[i] = -1;
    /* face maps per orientation */
    int64_t NFX = cx1 * ny * nz, NFY = (int64_t)nx * cy1 * nz,
            NFZ = (int64_t)nx * ny * cz1;
    uint8_t* fx = (uint8_t*)calloc((size_t)NFX, 1);
    uint8_t* fy = (uint8_t*)calloc((size_t)NFY, 1);
    uint8_t* fz = (uint8_t*)calloc((size_t)NFZ, 1);
    uint8_t* thin = eu ? (uint8_t*)calloc((size_t)N, 1) : NULL;
    int64_t cap_v = 1024, cap_q = 1024, nv = 0, nq = 0;
    int64_t* vcorner = (int64_t*)malloc(sizeof(int64_t) * cap_v);
    float* vcoord = (float*)malloc(sizeof(float) * 3 * cap_v);
    int32_t* qv = (int32_t*)malloc(sizeof(int32_t) * 4 * cap_q);

    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k)
            {
                int cur = occ[IDX(i, j, k)] != 0;
                for (int o = 0; o < 6; ++o)
                {
                    int a = i + S_OFF[o][0], b = j + S_OFF[o][1],
                        c = k + S_OFF[o][2];
                    int nb = occ_at(occ, nx, ny, nz, a, b, c);
                    if (cur == nb)
                        continue;
                    if (thin)
                    { /* save_to_thin_inside_layer (:597-617) */
                        int ux = cur ? i : a, uy = cur ? j : b,
                            uz = cur ? k : c;
                        for (int dx = -1; dx <= 1; ++dx)
                            for (int dy = -1; dy <= 1; ++dy)
                                for (int dz = -1; dz <= 1; ++dz)
                                    if (occ_at(occ, nx, ny, nz, ux + dx,
                                               uy + dy, uz + dz))
                                        thin[IDX(ux + dx, uy + dy, uz + dz)] =
                                            1;
                    }
                    int64_t cidx[8];
                    for (int cc = 0; cc < 8; ++cc)
                    {
                        int lx = i + (S_CSGN[cc][0] > 0),
                            ly = j + (S_CSGN[cc][1] > 0),
                            lz = k + (S_CSGN[cc][2] > 0);
                        cidx[cc] = ((int64_t)lx * cy1 + ly) * cz1 + lz;
                    }
                    for (int ii = 0; ii < 4; ++ii)
                    {
                        int ci = S_CWRT[o][ii];
                        if (cmap[cidx[ci]] >= 0)
                            continue;
                        if (nv == cap_v)
                        {
                            cap_v *= 2;
                            vcorner = (int64_t*)realloc(
                                vcorner, sizeof(int64_t) * cap_v);
                            vcoord = (float*)realloc(vcoord,
                                                     sizeof(float) * 3 * cap_v);
                        }
                        vcorner[nv] = cidx[ci];
                        vcoord[3 * nv + 0] = (float)i + 0.5f * S_CSGN[ci][0];
                        vcoord[3 * nv + 1] = (float)j + 0.5f * S_CSGN[ci][1];
                        vcoord[3 * nv + 2] = (float)k + 0.5f * S_CSGN[ci][2];
                        cmap[cidx[ci]] = (int32_t)nv;
                        nv++;
                    }
                    int fi = S_FWRT[o];
                    uint8_t* slot;
                    switch (fi)
                    {
                    case 0: slot = &fz[IDX(i, j, 0) / nz * cz1 + k]; break;
                    case 1: slot = &fz[IDX(i, j, 0) / nz * cz1 + k + 1]; break;
                    case 2: slot = &fx[((int64_t)(i + 1) * ny + j) * nz + k]; break;
                    case 3: slot = &fx[((int64_t)i * ny + j) * nz + k]; break;
                    case 4: slot = &fy[((int64_t)i * cy1 + j) * nz + k]; break;
                    default: slot = &fy[((int64_t)i * cy1 + j + 1) * nz + k]; break;
                    }
                    if (*slot)
                        continue;
                    *slot = 1;
                    if (nq == cap_q)
                    {
                        cap_q *= 2;
                        qv = (int32_t*)realloc(qv, sizeof(int32_t) * 4 * cap_q);
                    }
                    for (int t = 0; t < 4; ++t)
                        qv[4 * nq + t] = cmap[cidx[S_FCRN[fi][t]]];
                    nq++;
                }
            }

    if (eu)
    {
        int32_t tbl[256];
        or_cube_components(tbl);
        memset(eu, 0, sizeof(*eu));
        /* boundary vertices: 2x2x2 mask at (int)-truncated base voxel
         * (include/spaceinfo.h:140-153), bit (7-i) for coordmap[i] */
        int64_t V = 0;
        for (int64_t v = 0; v < nv; ++v)
        {
            int bx = (int)vcoord[3 * v + 0], by = (int)vcoord[3 * v + 1],
                bz = (int)vcoord[3 * v + 2];
            int mask = 0;
            for (int ci = 0; ci < 8; ++ci)
                mask |= occ_at(occ, nx, ny, nz, bx + ((ci >> 2) & 1),
                               by + ((ci >> 1) & 1), bz + (ci & 1))
                        << (7 - ci);
            V += tbl[mask];
        }
        /* boundary edges: multiplicity 1 if shared by 2 quads, else 2 */
        uint8_t* ecnt[3];
        for (int a = 0; a < 3; ++a)
            ecnt[a] = (uint8_t*)calloc((size_t)NC, 1);
        for (int64_t q = 0; q < nq; ++q)
            for (int t = 0; t < 4; ++t)
            {
                int64_t c0 = vcorner[qv[4 * q + t]];
                int64_t c1 = vcorner[qv[4 * q + (t + 1) % 4]];
                int64_t lo = c0 < c1 ? c0 : c1, hi = c0 < c1 ? c1 : c0;
                int64_t d = hi - lo;
                int axis = d == 1 ? 2 : (d == cz1 ? 1 : 0);
                ecnt[axis][lo]++;
            }
        int64_t E = 0;
        for (int a = 0; a < 3; ++a)
        {
            for (int64_t c = 0; c < NC; ++c)
                if (ecnt[a][c])
                    E += ecnt[a][c] == 2 ? 1 : 2;
            free(ecnt[a]);
        }
        /* interior counts (:394-427) */
        int64_t Vi = 0, Ei = 0, Fi = 0, T = 0;
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j)
                for (int k = 0; k < nz; ++k)
                {
                    if (!occ[IDX(i, j, k)])
                        continue;
                    T++;
                    for (int cc = 0; cc < 8; ++cc)
                    { /* vox_corner_offset (src/spaceinfo.cpp:3-8) */
                        float px = i + ((cc >> 2) & 1 ? 0.5f : -0.5f);
                        float py = j + ((cc >> 1) & 1 ? 0.5f : -0.5f);
                        float pz = k + (cc & 1 ? 0.5f : -0.5f);
                        int bx = (int)px, by = (int)py, bz = (int)pz;
                        int cnt = 0;
                        for (int ci = 0; ci < 8; ++ci)
                            cnt += occ_at(occ, nx, ny, nz, bx + ((ci >> 2) & 1),
                                          by + ((ci >> 1) & 1), bz + (ci & 1));
                        Vi += cnt == 8;
                    }
                    for (int ax = 0; ax < 3; ++ax)
                        for (int r1 = -1; r1 <= 0; ++r1)
                            for (int r2 = -1; r2 <= 0; ++r2)
                            {
                                int cnt = 0;
                                for (int u = r1; u <= r1 + 1; ++u)
                                    for (int w = r2; w <= r2 + 1; ++w)
                                    {
                                        int o3[3] = {0, 0, 0};
                                        o3[(ax + 1) % 3] = u;
                                        o3[(ax + 2) % 3] = w;
                                        cnt += occ_at(occ, nx, ny, nz, i + o3[0],
                                                      j + o3[1], k + o3[2]);
                                    }
                                Ei += cnt == 4;
                            }
                    for (int o = 0; o < 6; ++o)
                        Fi += occ_at(occ, nx, ny, nz, i + S_OFF[o][0],
                                     j + S_OFF[o][1], k + S_OFF[o][2]);
                }
        Vi /= 8;
        Ei /= 4;
        Fi /= 2;
        /* C: 6-components of the thin inside layer (:619-660) */
        int64_t C = padded_components(thin, nx, ny, nz, 1, 6, 0);
        eu->V_bnd = V;
        eu->E_bnd = E;
        eu->V_int = Vi;
        eu->E_int = Ei;
        eu->F_int = Fi;
        eu->V = V + Vi;
        eu->E = E + Ei;
        eu->F = nq + Fi;
        eu->T = T;
        eu->C = C;
        eu->euler = (int64_t)(int32_t)(V + Vi - E - Ei + nq + Fi - T);
    }

    int rc = 0;
    *n_sites = nv;
    *n_quads = nq;
    if (sites)
    {
        if (nv > site_cap)
            rc = -1;
        else
            memcpy(sites, vcoord, sizeof(float) * 3 * nv);
    }
    if (quads)
    {
        if (nq > quad_cap)
            rc = -1;
        else
            memcpy(quads, qv, sizeof(int32_t) * 4 * nq);
    }
    free(cmap);
    free(fx);
    free(fy);
    free(fz);
    free(thin);
    free(vcorner);
    free(vcoord);
    free(qv);
    return rc;
}

/* ======================================================================
 * Cell complex + FIFO simple-pair collapse (reference-pinned, R10-R12)
 * ==================================================================== */
typedef struct
{
    int64_t* keys; /* (a << 32) | (uint32)b, -1 = empty */
    int32_t* vals;
    int64_t cap;
} pair_map;

static uint64_t mix64(uint64_t x)
{
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    return x ^ (x >> 33);
}

static void pm_init(pair_map* m, int64_t n)
{
    int64_t cap = 16;
    while (cap < 2 * n + 16)
        cap <<= 1;
    m->cap = cap;
    m->keys = (int64_t*)malloc(sizeof(int64_t) * cap);
    m->vals = (int32_t*)malloc(sizeof(int32_t) * cap);
    for (int64_t i = 0; i < cap; ++i)
        m->keys[i] = -1;
}

/* returns the slot of key (existing or empty) */
static int64_t pm_slot(const pair_map* m, int64_t key)
{
    int64_t h = (int64_t)(mix64((uint64_t)key) & (uint64_t)(m->cap - 1));
    while (m->keys[h] != -1 && m->keys[h] != key)
        h = (h + 1) & (m->cap - 1);
    return h;
}

static int64_t pair_key(int32_t a, int32_t b)
{
    return ((int64_t)(uint32_t)a << 32) | (int64_t)(uint32_t)b;
}

int or_cc_finalize(int nv, int ne_in, const int32_t* edges_in, int nf,
                   const int64_t* foff, const int32_t* fv, int32_t* edges_out,
                   int edge_cap, int32_t* fe_out)
{
    pair_map m;
    pm_init(&m, (int64_t)ne_in + (nf ? foff[nf] : 0));
    int ne = 0;
    /* existing edges keep their order; duplicates collapse */
    for (int e = 0; e < ne_in; ++e)
    {
        int64_t k = pair_key(edges_in[2 * e], edges_in[2 * e + 1]);
        int64_t s = pm_slot(&m, k);
        if (m.keys[s] == -1)
        {
            m.keys[s] = k;
            m.vals[s] = ne;
            if (ne < edge_cap)
            {
                edges_out[2 * ne] = edges_in[2 * e];
                edges_out[2 * ne + 1] = edges_in[2 * e + 1];
            }
            ne++;
        }
    }
    /* face sides (makeEdge-normalised) in face order */
    for (int f = 0; f < nf; ++f)
    {
        int64_t n = foff[f + 1] - foff[f];
        for (int64_t i = 0; i < n; ++i)
        {
            int32_t u = fv[foff[f] + i], w = fv[foff[f] + (i + 1) % n];
            int32_t a = u < w ? u : w, b = u < w ? w : u;
            int64_t k = pair_key(a, b);
            int64_t s = pm_slot(&m, k);
            if (m.keys[s] == -1)
            {
                m.keys[s] = k;
                m.vals[s] = ne;
                if (ne < edge_cap)
                {
                    edges_out[2 * ne] = a;
                    edges_out[2 * ne + 1] = b;
                }
                ne++;
            }
            fe_out[foff[f] + i] = m.vals[s];
        }
    }
    free(m.keys);
    free(m.vals);
    (void)nv;
    return ne <= edge_cap ? ne : -1;
}

typedef struct
{
    int type; /* 1 = face-edge, 0 = edge-vertex */
    int32_t a, b;
} or_pair;

int or_cc_prune(int nv, int ne, const int32_t* edges, int nf,
                const int64_t* foff, const int32_t* fe, const float* em,
                const float* fm, float f_t, float l_t, uint8_t* rm_v,
                uint8_t* rm_e, uint8_t* rm_f, int32_t* ref_e)
{
    /* adjacency: vertex -> edges in edge order, edge -> faces in face order
     * (src/cellcomplex.cpp:456-474) */
    int64_t* voff = (int64_t*)calloc((size_t)nv + 1, sizeof(int64_t));
    int64_t* eoff = (int64_t*)calloc((size_t)ne + 1, sizeof(int64_t));
    for (int e = 0; e < ne; ++e)
    {
        voff[edges[2 * e] + 1]++;
        voff[edges[2 * e + 1] + 1]++;
    }
    for (int64_t i = 0; i < (nf ? foff[nf] : 0); ++i)
        eoff[fe[i] + 1]++;
    for (int v = 0; v < nv; ++v)
        voff[v + 1] += voff[v];
    for (int e = 0; e < ne; ++e)
        eoff[e + 1] += eoff[e];
    int32_t* ve = (int32_t*)malloc(sizeof(int32_t) * (size_t)(voff[nv] + 1));
    int32_t* ef = (int32_t*)malloc(sizeof(int32_t) * (size_t)(eoff[ne] + 1));
    int64_t* vfill = (int64_t*)calloc((size_t)nv + 1, sizeof(int64_t));
    int64_t* efill = (int64_t*)calloc((size_t)ne + 1, sizeof(int64_t));
    for (int e = 0; e < ne; ++e)
    {
        int a = edges[2 * e], b = edges[2 * e + 1];
        ve[voff[a] + vfill[a]++] = e;
        ve[voff[b] + vfill[b]++] = e;
    }
    for (int f = 0; f < nf; ++f)
        for (int64_t i = foff[f]; i < foff[f + 1]; ++i)
        {
            int e = fe[i];
            ef[eoff[e] + efill[e]++] = f;
        }
    int32_t* rv = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nv + 1));
    for (int v = 0; v < nv; ++v)
        rv[v] = (int32_t)(voff[v + 1] - voff[v]);
    for (int e = 0; e < ne; ++e)
        ref_e[e] = (int32_t)(eoff[e + 1] - eoff[e]);
    memset(rm_v, 0, (size_t)nv);
    memset(rm_e, 0, (size_t)ne);
    memset(rm_f, 0, (size_t)nf);

    int64_t qcap = 1024, qh = 0, qt = 0;
    or_pair* q = (or_pair*)malloc(sizeof(or_pair) * qcap);
#define QPUSH(T, A, B)                                                         \
    do                                                                         \
    {                                                                          \
        if (qt == qcap)                                                        \
        {                                                                      \
            qcap *= 2;                                                         \
            q = (or_pair*)realloc(q, sizeof(or_pair) * qcap);                  \
        }                                                                      \
        q[qt].type = (T);                                                      \
        q[qt].a = (A);                                                         \
        q[qt].b = (B);                                                         \
        qt++;                                                                  \
    } while (0)

    for (int e = 0; e < ne; ++e)
        if (ref_e[e] == 1)
        {
            int f = ef[eoff[e]];
            if (fm[f] < f_t)
                QPUSH(1, f, e);
        }
    for (int v = 0; v < nv; ++v)
        if (rv[v] == 1)
        {
            int e = ve[voff[v]];
            if (em[e] < l_t)
                QPUSH(0, e, v);
        }
    int rc = 0;
    while (qh < qt)
    {
        or_pair p = q[qh++];
        if (p.type == 1 ? rm_f[p.a] : rm_e[p.a])
            continue;
        if (!((p.type == 1 && ref_e[p.b] == 1) ||
              (p.type == 0 && rv[p.b] == 1)))
        {
            rc = -1; /* the reference prints "logic err" and exits */
            break;
        }
        if (p.type == 0)
        {
            rm_e[p.a] = 1;
            rm_v[p.b] = 1;
            int a = edges[2 * p.a], b = edges[2 * p.a + 1];
            rv[a]--;
            rv[b]--;
            int other = a != p.b ? a : b;
            if (rv[other] == 1)
                for (int64_t i = voff[other]; i < voff[other + 1]; ++i)
                {
                    int e2 = ve[i];
                    if (!rm_e[e2])
                    {
                        if (em[e2] < l_t)
                            QPUSH(0, e2, other);
                        break;
                    }
                }
        }
        else
        {
            rm_f[p.a] = 1;
            rm_e[p.b] = 1;
            ref_e[p.b]--;
            for (int64_t i = foff[p.a]; i < foff[p.a + 1]; ++i)
            {
                int e2 = fe[i];
                if (rm_e[e2])
                    continue;
                ref_e[e2]--;
                if (ref_e[e2] == 1)
                    for (int64_t j = eoff[e2]; j < eoff[e2 + 1]; ++j)
                    {
                        int f2 = ef[j];
                        if (!rm_f[f2])
                        {
                            if (fm[f2] < f_t)
                                QPUSH(1, f2, e2);
                            break;
                        }
                    }
            }
            for (int k = 0; k < 2; ++k)
            {
                int vi = edges[2 * p.b + k];
                rv[vi]--;
                if (rv[vi] == 1)
                    for (int64_t i = voff[vi]; i < voff[vi + 1]; ++i)
                    {
                        int e2 = ve[i];
                        if (!rm_e[e2])
                        {
                            if (em[e2] < l_t)
                                QPUSH(0, e2, vi);
                            break;
                        }
                    }
            }
        }
    }
#undef QPUSH
    free(q);
    free(voff);
    free(eoff);
    free(ve);
    free(ef);
    free(vfill);
    free(efill);
    free(rv);
    return rc;
}

int or_simple_triangulate(const int32_t* poly, int n, int32_t* tris)
{
    int i = n - 1, j = 0, t = 0, inc_j = 1;
    while (i - 1 > j)
    {
        if (inc_j)
        {
            tris[3 * t] = poly[i];
            tris[3 * t + 1] = poly[j];
            tris[3 * t + 2] = poly[j + 1];
            j++;
        }
        else
        {
            tris[3 * t] = poly[i];
            tris[3 * t + 1] = poly[j];
            tris[3 * t + 2] = poly[i - 1];
            i--;
        }
        t++;
        inc_j = !inc_j;
    }
    return t;
}
