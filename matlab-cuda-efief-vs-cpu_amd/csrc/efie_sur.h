// efie_sur.h -- host-side pieces of the near-pair surrogate shared by the
// context code (efie_api.cpp) and the trainer (efie_surrogate.cpp).
#pragma once

#include "../../include/efie.h"
#include "../../include/efie_surrogate.h"
#include "efie_internal.h"

#include <hip/hip_runtime.h>

#include <vector>

namespace efie {

// Model blob (floats): device model (parameters, then xmean[13],
// xinv[13] = 1/std, ymean[16], ystd[16]), then xmin[13], xmax[13] and the
// metadata below.
constexpr int kSurModel = kMlpParams + kMlpNorm;
constexpr int kSurBlob = kSurModel + 2 * kMlpIn + 8;
enum {
    kMetaVersion = 0,
    kMetaNearAlpha,
    kMetaNearSubdiv,
    kMetaNearOrder,
    kMetaQuadOrder,
    kMetaKlMax,
    kMetaDlMin,
    kMetaDlMax
};
constexpr float kSurVersion = 1.0f;
// centroid distance over L accepted by the surrogate
constexpr double kSurDlMin = 0.2, kSurDlMax = 3.0;

// Frame of the pair (t1, t2) from the vertices of t1 (v1, 3 x xyz) and t2
// (v2): fills p (except t1, t2) and, when non-null, tri18 (the six
// vertices in normalised frame coordinates); returns d / L.
double sur_frame(const double* v1, const double* v2, SurPair* p, double* tri18);

// Rule tables of the label kernel: points of the near_subdiv-times
// subdivided near_order rule and of the quad_order rule, as barycentric
// weights of the parent triangle's vertices plus the weight as a fraction
// of twice the parent's area ([n][4] each).
void sur_rule_tables(int quad_order, int near_order, int near_subdiv, std::vector<double>& sub,
                     std::vector<double>& reg);

// Records an error message for efie_last_error(); returns st.
efie_status sur_fail(efie_status st, const char* id, const char* msg);

// Samples, labels and trains on the current device (stream s); blob gets
// kSurBlob floats.
efie_status sur_train(hipStream_t s, const efie_opts* o, const efie_sur_cfg* cfg, std::vector<float>& blob,
                      efie_sur_report* rep);

}  // namespace efie
