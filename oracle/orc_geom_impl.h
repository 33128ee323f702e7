/*
 * ORACLE (test infrastructure only): precision-generic geometry of fgrain's hot path.
 *
 * Included twice by oracle.c, once with R=double (the reference restatement, FP64 as
 * in the Rust code) and once with R=float (the FP32 twin whose operation order is the
 * specification the GPU kernels follow, used for bit-exact first-hit comparisons).
 * Compiled with -ffp-contract=off so every product is rounded before every sum.
 *
 * Macros expected: R (float|double), FN(name) suffixing, RMIN/RMAX (NaN-ignoring
 * min/max like Rust's f64::min/max), RSQRT (sqrt), REXP.
 */

/* nalgebra UnitQuaternion::inverse_transform_vector: rotate by the conjugate (w, -u):
 * t = 2 (u' x v), out = (t * w + u' x t) + v with u' = -u  (crystal.rs:151-154). */
static inline void FN(qinv_rotate)(const R q[4], const R v[3], R out[3])
{
    const R w = q[0];
    const R ux = -q[1], uy = -q[2], uz = -q[3];
    const R tx = (uy * v[2] - uz * v[1]) * (R)2;
    const R ty = (uz * v[0] - ux * v[2]) * (R)2;
    const R tz = (ux * v[1] - uy * v[0]) * (R)2;
    const R cx = uy * tz - uz * ty;
    const R cy = uz * tx - ux * tz;
    const R cz = ux * ty - uy * tx;
    out[0] = (tx * w + cx) + v[0];
    out[1] = (ty * w + cy) + v[1];
    out[2] = (tz * w + cz) + v[2];
}

/* HalideCrystal::intersect_ray (crystal.rs:140-217).  dims are the already-expanded
 * semi-axes (dimensions * expansion).  Returns 1 and (t0, t1) on a hit. */
static int FN(crystal_intersect)(const R c[3], const R q[4], const R dims[3], int morph,
                                 const R o[3], const R d[3], R *t0, R *t1)
{
    R rel[3] = {o[0] - c[0], o[1] - c[1], o[2] - c[2]};
    R lo[3], ld[3];
    FN(qinv_rotate)(q, rel, lo);
    FN(qinv_rotate)(q, d, ld);
    if (morph == 0) {
        /* cubic: oriented slab test with +-inf reciprocals (crystal.rs:160-179) */
        const R ix = (R)1 / ld[0], iy = (R)1 / ld[1], iz = (R)1 / ld[2];
        const R a1 = (-dims[0] - lo[0]) * ix, a2 = (dims[0] - lo[0]) * ix;
        const R a3 = (-dims[1] - lo[1]) * iy, a4 = (dims[1] - lo[1]) * iy;
        const R a5 = (-dims[2] - lo[2]) * iz, a6 = (dims[2] - lo[2]) * iz;
        const R tmin = RMAX(RMAX(RMIN(a1, a2), RMIN(a3, a4)), RMIN(a5, a6));
        const R tmax = RMIN(RMIN(RMAX(a1, a2), RMAX(a3, a4)), RMAX(a5, a6));
        if (tmax >= tmin && tmax > (R)0) {
            *t0 = RMAX(tmin, (R)0);
            *t1 = tmax;
            return 1;
        }
        return 0;
    }
    /* tabular / octahedral: oriented ellipsoid (crystal.rs:180-215) */
    const R px = lo[0] / dims[0], py = lo[1] / dims[1], pz = lo[2] / dims[2];
    const R dx = ld[0] / dims[0], dy = ld[1] / dims[1], dz = ld[2] / dims[2];
    const R a = dx * dx + dy * dy + dz * dz;
    const R b = (R)2 * (px * dx + py * dy + pz * dz);
    const R cc = (px * px + py * py + pz * pz) - (R)1;
    const R disc = b * b - (R)4 * a * cc;
    if (disc < (R)0) return 0;
    const R s = RSQRT(disc);
    const R r0 = (-b - s) / ((R)2 * a);
    const R r1 = (-b + s) / ((R)2 * a);
    const R te = RMIN(r0, r1), tx = RMAX(r0, r1);
    if (tx > (R)0) {
        *t0 = RMAX(te, (R)0);
        *t1 = tx;
        return 1;
    }
    return 0;
}

/* AABB::intersect_ray (bvh.rs:96-124), Kay-Kajiya with precomputed inverse direction. */
static inline int FN(box_hit)(const R mn[3], const R mx[3], const R o[3], const R inv[3])
{
    const R t1x = (mn[0] - o[0]) * inv[0], t2x = (mx[0] - o[0]) * inv[0];
    const R tminx = RMIN(t1x, t2x), tmaxx = RMAX(t1x, t2x);
    const R t1y = (mn[1] - o[1]) * inv[1], t2y = (mx[1] - o[1]) * inv[1];
    const R tminy = RMIN(t1y, t2y), tmaxy = RMAX(t1y, t2y);
    const R t1z = (mn[2] - o[2]) * inv[2], t2z = (mx[2] - o[2]) * inv[2];
    const R tminz = RMIN(t1z, t2z), tmaxz = RMAX(t1z, t2z);
    const R te = RMAX(RMAX(tminx, tminy), tminz);
    const R tx = RMIN(RMIN(tmaxx, tmaxy), tmaxz);
    return (tx >= te && tx > (R)0);
}
