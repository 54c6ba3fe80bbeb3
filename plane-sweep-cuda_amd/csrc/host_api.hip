// host_api.hip -- host-side pieces of the reference's drivers that callers of
// the kernels need: the float plane-depth loop, RelativeMatrices, and the
// per-plane homography in the reference's float Matrix3D arithmetic.  Pure
// host code (built with -ffp-contract=off like the kernels), so a caller that
// feeds these into psw_sweep_fused_reference reproduces
// PlaneSweep::RunAlgorithm's float results.
#include "psw_common.h"

namespace psw {

// Matrix3D::det / ::inv (src/inc/structs.h:565-593), float.
static M3f inv3f(const M3f& M) {
    const float* m = M.m;
    const float det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                      m[2] * (m[3] * m[7] - m[4] * m[6]);
    const float r[9] = {m[4] * m[8] - m[5] * m[7], m[2] * m[7] - m[1] * m[8], m[1] * m[5] - m[2] * m[4],
                        m[5] * m[6] - m[3] * m[8], m[0] * m[8] - m[2] * m[6], m[2] * m[3] - m[0] * m[5],
                        m[3] * m[7] - m[4] * m[6], m[1] * m[6] - m[0] * m[7], m[0] * m[4] - m[1] * m[3]};
    M3f o;
    for (int j = 0; j < 9; ++j) o.m[j] = r[j] / det;
    return o;
}

static M3f trans3f(const M3f& M) {
    M3f o;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) o.m[3 * j + i] = M.m[3 * i + j];
    return o;
}

// PlaneSweep::RelativeMatrices (src/planesweep.cpp:538-549), float.
static void relative3f(const float* Rref, const float* tref, const float* Rsrc, const float* tsrc, bool alt,
                       M3f& Rrel, V3f& trel) {
    const M3f Rr = m3_from(Rref), Rs = m3_from(Rsrc);
    if (!alt) {
        Rrel = mm3(Rs, inv3f(Rr));
        float a, b, c;
        mv3(Rrel, tref[0], tref[1], tref[2], a, b, c);
        trel = V3f{tsrc[0] - a, tsrc[1] - b, tsrc[2] - c};
    } else {
        const M3f RsT = trans3f(Rs);
        Rrel = mm3(RsT, Rr);
        float a, b, c;
        mv3(RsT, tref[0] - tsrc[0], tref[1] - tsrc[1], tref[2] - tsrc[2], a, b, c);
        trel = V3f{a, b, c};
    }
}

}  // namespace psw

using namespace psw;

extern "C" {

// The reference's plane loop: dstep = (zfar - znear) / (nplanes - 1) and
// for (float d = znear; d <= zfar; d += dstep) (src/planesweep.cpp:169, :200).
int psw_reference_planes(float znear, float zfar, unsigned nplanes, float* out, int maxn, int* count) {
    PSW_CHECK_ARG(out && count && maxn >= 0, "NULL pointer");
    const float dstep = (zfar - znear) / (float)(nplanes - 1);
    int n = 0;
    for (float d = znear; d <= zfar; d += dstep) {
        if (n >= maxn) break;
        out[n++] = d;
    }
    *count = n;
    return PSW_OK;
}

int psw_mat3_inverse(const float* m, float* out) {
    PSW_CHECK_ARG(m && out, "NULL pointer");
    const M3f r = inv3f(m3_from(m));
    for (int j = 0; j < 9; ++j) out[j] = r.m[j];
    return PSW_OK;
}

int psw_relative_matrices(const float* Rref, const float* tref, const float* Rsrc, const float* tsrc, int alternative,
                          float* Rrel, float* trel) {
    PSW_CHECK_ARG(Rref && tref && Rsrc && tsrc && Rrel && trel, "NULL pointer");
    M3f R;
    V3f t;
    relative3f(Rref, tref, Rsrc, tsrc, alternative != 0, R, t);
    for (int j = 0; j < 9; ++j) Rrel[j] = R.m[j];
    trel[0] = t.x; trel[1] = t.y; trel[2] = t.z;
    return PSW_OK;
}

// H = K * (Rrel + tr / d) * invK; H = H / H(2,2) with tr = trel e3^T, all in
// float Matrix3D arithmetic (src/planesweep.cpp:195-203).  Output host array
// H[nsrc][nplanes][9].
int psw_sweep_homographies_reference(float* H, const float* depths, int nplanes, int nsrc, const float* K,
                                     const float* Rref, const float* tref, const float* Rsrc, const float* tsrc,
                                     int alternative) {
    PSW_CHECK_ARG(H && depths && K && Rref && tref && Rsrc && tsrc, "NULL pointer");
    PSW_CHECK_ARG(nplanes >= 1 && nsrc >= 1, "need planes and sources");
    const M3f Km = m3_from(K), invK = inv3f(Km);
    for (int s = 0; s < nsrc; ++s) {
        M3f Rrel;
        V3f trel;
        relative3f(Rref, tref, Rsrc + 9 * s, tsrc + 3 * s, alternative != 0, Rrel, trel);
        M3f tr = {{0.f, 0.f, trel.x, 0.f, 0.f, trel.y, 0.f, 0.f, trel.z}};
        for (int p = 0; p < nplanes; ++p) {
            const float d = depths[p];
            M3f M;
            for (int j = 0; j < 9; ++j) M.m[j] = Rrel.m[j] + tr.m[j] / d;
            const M3f Hm = mm3(mm3(Km, M), invK);
            const float h22 = Hm.m[8];
            float* o = H + ((size_t)s * nplanes + p) * 9;
            for (int j = 0; j < 9; ++j) o[j] = Hm.m[j] / h22;
        }
    }
    return PSW_OK;
}

}  // extern "C"
