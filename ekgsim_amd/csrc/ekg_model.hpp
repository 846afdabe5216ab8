// ekg_model.hpp -- host FP64 model functions around the GPU simulation:
// WohlfartPlus values, parameter vector -> layer APs, ECG comparison.
#pragma once

#include <cstddef>

#include "ekg.h"

namespace ekg {

typedef ekg_ap_model ApModel;

// WohlfartPlus W(t) with coefficients k[0..8] (Wohlfart.h:195-203).
double ap_value(const double* k, double t);
// Time of 90 % repolarisation on the 1 ms grid (Wohlfart.h:206-223), -1 if none.
double ap_apd90(const double* k);

// Entries of the parameter vector that the layer APs consume.
int ap_num_genes(const ApModel& m);
// Layer APs [n_layers][9] from the parameter vector; violation is the summed
// distance of free coefficients outside [k min, k max].
void layer_aps(const ApModel& m, int n_layers, const double* genes, int n_genes, double* layer_k,
               double* violation);

// Comparison value of a simulated ECG a against a target b (EKG_CMP_*); the
// correlation modes return 1 - r.
double criterion(int mode, const double* a, size_t na, const double* b, size_t nb, double b_offset);

// Linear resampling of a target curve; returns the output length and writes
// at most cap samples.
size_t resample(const double* src, size_t n, double factor, int start, double* dest, size_t cap);

}  // namespace ekg
