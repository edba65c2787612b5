// y3g_internal.hpp -- argument blocks shared by the kernels
// (y3g_kernels.hip) and the host runtime (y3g_runtime.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "y3g.h"
#include "y3g_device.hpp"

namespace y3g {

struct TableArgs {
  const PointDev* points;
  y3g_table_spec spec;
  double* hmf64;
  float* hmf32;
  int hmf_stride;
  double* da64;
  float* da32;
  int da_stride;
  double* sci64;
  float* sci32;
  int sci_stride;
  // the HMF table's cubic-spline coefficients (hmf_coefs), (n_z + 4) x
  // (n_lnm + 4) per point
  double* hmfc64;
  float* hmfc32;
  int hmfc_stride;
};

struct KernArgs {
  const y3g_integral* igs;
  int n;
  const int* order;
  const PointDev* points;
  const float* hmf32;
  const double* hmf64;
  int hmf_stride;
  const float* da32;
  const double* da64;
  int da_stride;
  const float* sci32;
  const double* sci64;
  int sci_stride;
  // the HMF coefficients; null: the synthetic integrands read the bilinear
  // HMF (y3g_options.hmf_smooth = 0)
  const float* hmfc32;
  const double* hmfc64;
  int hmfc_stride;
  // DES data of the points (Y3G_DSIGMA6D_DES), point p at des + des_off[p]
  const double* des;
  const long long* des_off;
  // cubature region slab: per slot lo/hi [kMaxDim][cap], est/err/axis [cap]
  double* slab_lo;
  double* slab_hi;
  double* slab_est;
  double* slab_err;
  double* slab_raw;  // the regions' own |R7-R5| (two-level estimate)
  int8_t* slab_ax;
  int cap;
  int kmax;
  int init_split;
  int init_lnm;   // initial pieces along the synthetic kinds' lnM axis (0: init_split)
  double err_scale;  // safety factor on the summed error with the smoothed HMF
  double ext_ratio;  // budget extension: error / tolerance above which ...
  int ext_regions;   // ... a spent budget is raised to this many regions (0: off)
  double err_c, err_smin;
  double tl_c, tl_smin;
  // VEGAS
  long long vegas_calls;
  int vegas_itmax, vegas_itskip, vegas_nbins;
  double vegas_alpha;
  int* queue;
  y3g_result* out;
  unsigned long long* nevals;
};

// Continuous batching (k_cubature_stream).  Ring entries and the control
// block live in coherent pinned host memory; StreamDev in device memory.
constexpr int kStreamJobIntegral = 0;
constexpr int kStreamJobTables = 1;
struct StreamItem {
  int job;           // kStreamJobIntegral or kStreamJobTables
  int slot;          // table slot (the integral's point)
  unsigned gen;      // the slot's table generation needed / built
  int pad;
  y3g_integral ig;   // integrals only
};
static_assert(sizeof(StreamItem) % 8 == 0, "ring entries are read as 64-bit words");
static_assert(sizeof(PointDev) % 8 == 0, "PointDev is read as 64-bit words");
struct StreamCtl {
  unsigned long long tail;  // entries published by the host
  unsigned close;           // host: exit once every published entry is claimed
  unsigned pad;
};
struct StreamDev {  // one 128-byte line per word that workgroups contend for
  alignas(128) unsigned long long head;  // entries claimed
  alignas(128) unsigned long long tail;  // the device's copy of StreamCtl::tail
  alignas(128) unsigned close;           // the device's copy of StreamCtl::close
  alignas(128) unsigned lock;            // held by the workgroup refreshing the copies
  alignas(128) unsigned busy;            // workgroups claiming or working on an entry
  alignas(128) long long last;           // wall clock of the last claim or finish
  alignas(128) unsigned quit;            // set once the device has been idle: all exit
};
struct StreamArgs {
  const StreamItem* ring;       // [ring_cap], host
  y3g_result* res;              // [ring_cap], host
  unsigned long long* seq;      // [ring_cap], host: entry + 1 once res is written
  const StreamCtl* ctl;         // host
  const PointDev* hpoints;      // [table slots], host: what a table build reads
  StreamDev* dev;               // device
  unsigned* ready;              // [table slots], device: generation of the tables
  PointDev* points;             // [table slots], device: the kernels' PointDev
  TableArgs tab;                // the slots' tables (written by table builds)
  unsigned ring_cap;
  long long idle_ticks, wait_ticks;  // wall-clock ticks
};

struct FunctorArgs {
  int which;
  const double* in;
  int nin;
  size_t n;
  const double* par;
  const PointDev* points;  // the point the table functors read
  const void* hmf;
  const void* da;
  const void* sci;
  const double* des;       // the point's DES data (Y3G_DSIGMA6D_DES integrands)
  const y3g_integral* ig;
  const double* nfwmis;    // NFW_DSIGMA_MIS table {ln x, ln xmis, T}
  int nfw_nx, nfw_nm;
  double* out;
};

// Fixed-GL counts with a selection function: per point, its selection
// function (axes and planes) at offsets into one FP64 buffer.
struct SelPoint {
  long long off_lnm, off_z, off_s;  // offsets into SelArgs::data
  int n_lnm, n_z, n_bins;
};
struct SelGLArgs {
  const PointDev* points;
  const double* hmf64;
  int hmf_stride;
  const double* da64;
  int da_stride;
  const SelPoint* sel;
  const double* data;
  const double* nodes;  // lnM nodes [n_lnm], weights [n_lnm], z nodes [n_z], weights [n_z]
  int n_lnm, n_z;
  int max_bins;
  int max_axes;         // largest n_lnm + n_z of the launch's selection functions
  double* out;          // [n_points][max_bins]
};
hipError_t launch_selgl(const SelGLArgs& A, int n_points, hipStream_t s);

size_t cubature_smem(int precision, int block);
// Workgroups of k_cubature resident per CU (registers, LDS), 0 on error.
int cubature_blocks_per_cu(int precision, int block);
hipError_t launch_cubature(const KernArgs& A, int precision, int block,
                           int slots, hipStream_t s);
hipError_t launch_vegas(const KernArgs& A, int precision, int block, int slots,
                        hipStream_t s);
// The streaming cubature (512-thread workgroups) and its residency per CU.
hipError_t launch_cubature_stream(const KernArgs& A, const StreamArgs& Z, int precision,
                                  int slots, hipStream_t s);
int cubature_stream_blocks_per_cu(int precision);
hipError_t launch_tables(const TableArgs& A, int n_points, hipStream_t s);
hipError_t launch_hmf_coefs(const TableArgs& A, int pt0, int n_points, hipStream_t s);
hipError_t launch_zero_words(void* p, size_t n, hipStream_t s);
hipError_t launch_nfw_mis_table(double* tab, int nx, int nm, hipStream_t s);
hipError_t launch_functor(const FunctorArgs& F, int precision, hipStream_t s);

}  // namespace y3g
