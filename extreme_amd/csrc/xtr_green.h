// Device assembly of the anomaly-to-anomaly Green tensor (FP64).
#pragma once

#include "xtr_kernels.h"

#include <cmath>
#include <vector>

namespace xtr {

// 1-D background section (Extreme.Core/Model/Section1D.cs): layer 0 the air half-space, the
// last layer the lower half-space, both thickness 0.
struct Section {
    double zero_level = 0;
    std::vector<double> thickness, sigma;
};

// The reference keeps depths and thicknesses as System.Decimal (Extreme.Core/Model/Layer1D.cs), so
// its layer lookups add and compare exactly; on doubles 2.2 + 99.9 lands above 102.1. The lookups
// below compare whole micrometres instead, exact for any depth written with at most six decimals.
inline long long depth_um(double v) { return std::llround(v * 1e6); }

// ModelUtils.FindCorrespondingBackgroundLayerIndex (Model/ModelUtils.cs:35-52): the first layer whose
// closed interval holds [start, end]; -1 when none does.
inline int background_layer_index(const Section &s, double start, double end)
{
    long long ss = depth_um(s.zero_level);
    const long long a = depth_um(start), b = depth_um(end);
    for (size_t i = 0; i < s.thickness.size(); ++i) {
        const long long es = ss + depth_um(s.thickness[i]);
        if (a >= ss && b <= es) return (int)i;
        ss = es;
    }
    return -1;
}

// ScalarMathUtils.DetermineCorrespondingBackgroundLayer (Green/Scalar/Impl/ScalarMathUtils.cs:15-31):
// the layer whose half-open interval [d1, d2) holds depth; -1 when none does.
inline int layer_index_at_depth(const Section &s, double depth)
{
    long long prev = depth_um(s.zero_level);
    const long long z = depth_um(depth);
    for (size_t k = 0; k < s.thickness.size(); ++k) {
        const long long d2 = prev + depth_um(s.thickness[k]);
        if (z >= prev && z < d2) return (int)k;
        prev = d2;
    }
    return -1;
}

struct GreenInput {
    Geom g;
    double dx = 0, dy = 0;
    std::vector<double> depth, thick;
    Section section;
    double omega = 0;
    int number_of_hankels = 10;
    std::vector<double> hank0, hank1;   // 431 taps each
};

// Writes the FFT'd tensor in the device layout [L][xx S | xy S | yy S | zz S | xz nz^2 | yz nz^2]
// into G (L * gblk complex128), ordered on stream s.
void build_green_atoa(const GreenInput &in, double2 *G, const double2 *tw_x, const double2 *tw_y, hipStream_t s);

// AtoO level tensor for the observation level at depth level_z whose sites are the cell centres
// shifted by (shift_x, shift_y) metres: device layout [L][component][nz] with components, in order,
// xx yy zz xy xz yz zx zy (electric = 1) or xx xy yx xz yz zx zy (electric = 0), FFT'd; G holds
// L * ncomp * nz complex128.
int atoo_components(int electric, int *slot /* [9]: xx yy zz xy yx xz yz zx zy -> compact or -1 */);
void build_green_atoo_level(const GreenInput &in, double level_z, double shift_x, double shift_y, int electric,
                            double2 *G, const double2 *tw_x, const double2 *tw_y, hipStream_t s);

}  // namespace xtr
