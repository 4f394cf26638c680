// Device assembly of the anomaly-to-anomaly Green tensor (FP64).
#pragma once

#include "xtr_kernels.h"

#include <vector>

namespace xtr {

// 1-D background section (Extreme.Core/Model/Section1D.cs): layer 0 the air half-space, the
// last layer the lower half-space, both thickness 0.
struct Section {
    double zero_level = 0;
    std::vector<double> thickness, sigma;
};

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
