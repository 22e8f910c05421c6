// efie_mex_common.h -- helpers shared by the MEX entry points.  Built with
// the interleaved-complex MATLAB API (mex -R2018a).
#pragma once

#include <string.h>

#include "efie.h"
#include "mex.h"

namespace efie_mex {

// One device context per MEX file, kept across calls and freed at unload.
inline efie_ctx*& ctx_slot() {
    static efie_ctx* c = nullptr;
    return c;
}
inline void release_ctx(void) {
    if (ctx_slot()) {
        efie_destroy(ctx_slot());
        ctx_slot() = nullptr;
    }
}

inline void check(bool cond, const char* id, const char* msg) {
    if (!cond) mexErrMsgIdAndTxt(id, msg);
}

// Raise the C ABI's error with the reference MEX identifier it carries.
inline void status(efie_status st) {
    if (st != EFIE_OK) mexErrMsgIdAndTxt(efie_last_error_id(), "%s", efie_last_error());
}

inline efie_ctx* context() {
    if (!ctx_slot()) {
        status(efie_create(&ctx_slot(), 0));
        mexAtExit(release_ctx);
    }
    return ctx_slot();
}

// complex scalar given as real or complex double (patch_sweep_cuda.cu:156-166)
inline void complex_scalar(const mxArray* a, double* re, double* im) {
    *re = mxGetScalar(a);
    *im = 0.0;
    if (mxIsComplex(a)) {
        const mxComplexDouble* z = mxGetComplexDoubles(a);
        *re = z[0].real;
        *im = z[0].imag;
    }
}

inline double field_or(const mxArray* s, const char* name, double def) {
    if (!s) return def;
    const mxArray* f = mxGetField(s, 0, name);
    if (!f || mxGetNumberOfElements(f) < 1) return def;
    return mxGetScalar(f);
}

// Optional trailing options struct with the reference opts field names
// (main_patch_efie_rwg.m:281-289) plus model / precision / batch.
inline void read_opts(const mxArray* s, efie_opts* o) {
    efie_opts_default(o);
    if (!s) return;
    check(mxIsStruct(s), "mex:opts", "opts must be a struct.");
    o->quad_order = (int)field_or(s, "quad_order", o->quad_order);
    o->near_sing_eps = field_or(s, "near_sing_eps", o->near_sing_eps);
    o->self_subdiv = (int)field_or(s, "self_subdiv", o->self_subdiv);
    o->self_alpha = field_or(s, "self_alpha", o->self_alpha);
    o->near_subdiv = (int)field_or(s, "near_subdiv", o->near_subdiv);
    o->near_alpha = field_or(s, "near_alpha", o->near_alpha);
    o->near_order = (int)field_or(s, "near_order", o->near_order);
    o->model = (int)field_or(s, "model", o->model);
    o->precision = (int)field_or(s, "precision", o->precision);
    o->batch = (int)field_or(s, "batch", o->batch);
    o->ir_tol = field_or(s, "ir_tol", o->ir_tol);
    o->ir_max_iter = (int)field_or(s, "ir_max_iter", o->ir_max_iter);
    o->near_mode = (int)field_or(s, "near_mode", o->near_mode);
    o->lu_kernel = (int)field_or(s, "lu_kernel", o->lu_kernel);
}

// Arguments 1-12 of both MEX functions: V, F, edge, plusTri, minusTri,
// plusSign, minusSign, rp, rm, Ap, Am, len, with the reference's checks
// (patch_sweep_cuda.cu:201-219).  edgeId is "mex:edge" or "mex:E".
// Runs before any device work.
inline void check_tables(const mxArray* const* prhs, const char* edgeId, const char* edgeMsg) {
    const mxArray *Vmx = prhs[0], *Fmx = prhs[1], *Emx = prhs[2];
    check(mxIsDouble(Vmx) && mxGetN(Vmx) == 3, "mex:V", "V must be [Nv x 3] double.");
    check(mxIsInt32(Fmx) && mxGetN(Fmx) == 3, "mex:F", "F must be [Nt x 3] int32.");
    check(mxIsInt32(Emx) && mxGetN(Emx) == 2, edgeId, edgeMsg);
    const mwSize Nv = mxGetM(Vmx), Nt = mxGetM(Fmx), Ne = mxGetM(Emx);
    check(mxGetM(prhs[3]) == Ne && mxIsInt32(prhs[3]), "mex:plusTri", "plusTri must be int32 [Ne x 1].");
    check(mxGetM(prhs[4]) == Ne && mxIsInt32(prhs[4]), "mex:minusTri", "minusTri must be int32 [Ne x 1].");
    check(mxGetM(prhs[5]) == Ne && mxIsInt32(prhs[5]), "mex:plusSign", "plusSign must be int32 [Ne x 1].");
    check(mxGetM(prhs[6]) == Ne && mxIsInt32(prhs[6]), "mex:minusSign", "minusSign must be int32 [Ne x 1].");
    check(mxIsDouble(prhs[7]) && mxGetM(prhs[7]) == Ne && mxGetN(prhs[7]) == 3, "mex:rp", "rp must be [Ne x 3] double.");
    check(mxIsDouble(prhs[8]) && mxGetM(prhs[8]) == Ne && mxGetN(prhs[8]) == 3, "mex:rm", "rm must be [Ne x 3] double.");
    check(mxIsDouble(prhs[9]) && mxGetNumberOfElements(prhs[9]) == Ne, "mex:Ap", "Ap must be [Ne x 1] double.");
    check(mxIsDouble(prhs[10]) && mxGetNumberOfElements(prhs[10]) == Ne, "mex:Am", "Am must be [Ne x 1] double.");
    check(mxIsDouble(prhs[11]) && mxGetNumberOfElements(prhs[11]) == Ne, "mex:len", "len must be [Ne x 1] double.");
    (void)Nv;
    (void)Nt;
}

// The prepack struct of efie_precompute_pack.m:1-59, with the checks of
// ggassemble_ZV_EFIE_RWG_DIFF_GPU_FAST.cu:608-641 plus the element counts
// that keep every read in bounds.  selfA names the field of the self
// softening lengths ("self_a"; the 8-input entry point also accepts
// "self_aSoft", the name its reference source reads).  The struct's
// near_alpha field is read when nearAlphaField is set.
inline void read_prepack(const mxArray* pk, mwSize Nt, bool nearAlphaField, efie_prepack* out) {
    check(pk && mxIsStruct(pk), "mex:prepack", "prepack must be a struct from efie_precompute_pack.");
    auto get = [&](const char* name) -> const mxArray* {
        const mxArray* f = mxGetField(pk, 0, name);
        check(f != nullptr, "mex:prepack", "Missing field in prepack.");
        return f;
    };
    const mxArray *rreg = get("r_reg"), *wreg = get("w_reg"), *nm = get("nearMat");
    const mxArray *sr = get("self_r"), *sw = get("self_w");
    const mxArray* sa = mxGetField(pk, 0, "self_a");
    if (!sa) sa = mxGetField(pk, 0, "self_aSoft");
    check(sa != nullptr, "mex:prepack", "Missing field in prepack.");
    const mxArray *nr = get("near_r"), *nw = get("near_w"), *aeq = get("aeqTri");
    check(mxIsDouble(rreg) && mxGetNumberOfDimensions(rreg) == 3, "mex:r_reg", "r_reg must be [Nq x 3 x Nt] double.");
    check(mxIsDouble(wreg) && mxGetNumberOfDimensions(wreg) == 2, "mex:w_reg", "w_reg must be [Nq x Nt] double.");
    check(mxIsUint8(nm), "mex:nearMat", "nearMat must be uint8 [Nt x Nt].");
    check(mxIsDouble(sr) && mxGetNumberOfDimensions(sr) == 4, "mex:self_r", "self_r must be [Nq2 x 3 x NsS x Nt].");
    check(mxIsDouble(sw) && mxGetNumberOfDimensions(sw) == 3, "mex:self_w", "self_w must be [Nq2 x NsS x Nt].");
    check(mxIsDouble(sa) && mxGetNumberOfDimensions(sa) == 2, "mex:self_a", "self_a must be [NsS x Nt].");
    check(mxIsDouble(nr) && mxGetNumberOfDimensions(nr) == 4, "mex:near_r", "near_r must be [Nq2 x 3 x NsN x Nt].");
    check(mxIsDouble(nw) && mxGetNumberOfDimensions(nw) == 3, "mex:near_w", "near_w must be [Nq2 x NsN x Nt].");
    check(mxIsDouble(aeq) && mxGetNumberOfElements(aeq) == Nt, "mex:aeqTri", "aeqTri must be [Nt x 1].");
    const mwSize* dr = mxGetDimensions(rreg);
    check(dr[1] == 3, "mex:r_reg", "r_reg dim2 must be 3.");
    check(dr[2] == Nt, "mex:r_reg", "r_reg dim3 must be Nt.");
    const mwSize Nq = dr[0];
    check(mxGetM(wreg) == Nq && mxGetN(wreg) == Nt, "mex:w_reg", "w_reg must be [Nq x Nt] double.");
    check(mxGetNumberOfElements(nm) == Nt * Nt, "mex:nearMat", "nearMat must be uint8 [Nt x Nt].");
    const mwSize* ds = mxGetDimensions(sr);
    check(ds[1] == 3, "mex:self_r", "self_r dim2 must be 3.");
    check(ds[3] == Nt, "mex:self_r", "self_r dim4 must be Nt.");
    const mwSize Nq2 = ds[0], NsS = ds[2];
    const mwSize* dn = mxGetDimensions(nr);
    check(dn[0] == Nq2, "mex:near_r", "near_r Nq2 mismatch.");
    check(dn[1] == 3, "mex:near_r", "near_r dim2 must be 3.");
    check(dn[3] == Nt, "mex:near_r", "near_r dim4 must be Nt.");
    const mwSize NsN = dn[2];
    check(mxGetNumberOfElements(sw) == Nq2 * NsS * Nt, "mex:self_w", "self_w must be [Nq2 x NsS x Nt].");
    check(mxGetNumberOfElements(sa) == NsS * Nt, "mex:self_a", "self_a must be [NsS x Nt].");
    check(mxGetNumberOfElements(nw) == Nq2 * NsN * Nt, "mex:near_w", "near_w must be [Nq2 x NsN x Nt].");
    out->Nq = (int)Nq;
    out->Nq2 = (int)Nq2;
    out->NsS = (int)NsS;
    out->NsN = (int)NsN;
    out->r_reg = mxGetDoubles(rreg);
    out->w_reg = mxGetDoubles(wreg);
    out->nearMat = mxGetUint8s(nm);
    out->self_r = mxGetDoubles(sr);
    out->self_w = mxGetDoubles(sw);
    out->self_a = mxGetDoubles(sa);
    out->near_r = mxGetDoubles(nr);
    out->near_w = mxGetDoubles(nw);
    out->aeqTri = mxGetDoubles(aeq);
    out->near_alpha = nearAlphaField ? mxGetScalar(get("near_alpha")) : 0.15;
}

inline void upload_tables(efie_ctx* c, const mxArray* const* prhs) {
    const mxArray *Vmx = prhs[0], *Fmx = prhs[1], *Emx = prhs[2];
    const mwSize Nv = mxGetM(Vmx), Nt = mxGetM(Fmx), Ne = mxGetM(Emx);
    status(efie_set_mesh(c, mxGetDoubles(Vmx), (int)Nv, mxGetInt32s(Fmx), (int)Nt));
    status(efie_set_rwg(c, mxGetInt32s(Emx), mxGetInt32s(prhs[3]), mxGetInt32s(prhs[4]), mxGetInt32s(prhs[5]),
                        mxGetInt32s(prhs[6]), mxGetDoubles(prhs[7]), mxGetDoubles(prhs[8]), mxGetDoubles(prhs[9]),
                        mxGetDoubles(prhs[10]), mxGetDoubles(prhs[11]), (int)Ne));
}

}  // namespace efie_mex
