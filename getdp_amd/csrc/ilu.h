// ILU(0) factor object shared by the solve driver and the Krylov methods.
#pragma once

#include "gdp_device.h"

namespace gdp {

struct IluFactor {
  int n = 0;
  long nnz = 0;
  const int *ia = nullptr, *ja = nullptr;  // pattern (not owned)
  int *diag = nullptr;                      // diagonal position per row
  float *lu32 = nullptr;
  double *lu64 = nullptr;
  bool is_f64 = false;
  int nlevels = 0, upper_nlevels = 0;
  int g_lower = 1, g_upper = 1;             // lanes per row in the two sweeps
  int maxrow = 1;                           // longest row of the pattern
  std::vector<int> level_ptr;               // host: rows of level l
  int *level_rows = nullptr;
  double *tmp64 = nullptr;                  // forward-sweep output (sentinel-filled)
  float *tmp32 = nullptr;
  int *err = nullptr;                       // [0] zero pivot, [1] timeout, [2] missing diag
  // factors computed on the host (ILUT / ILUTP): the object owns their pattern
  int *own_ia = nullptr, *own_ja = nullptr;
  // ILUTP: entry j of an application's output is entry colmap[j] of the
  // backward sweep's result, which is staged in pbuf64 / pbuf32
  int *colmap = nullptr;
  double *pbuf64 = nullptr;
  float *pbuf32 = nullptr;
  // One-wave sweeps, used when the dependency chain is long and the levels
  // are narrow (threshold ILU with fill): each sweep's rows in level order,
  // in chunks of up to 64 rows of one level, one lane per row; per-lane
  // fields at c * 64 + l.
  struct WaveSweep {
    int G = 1, wg = 0, nch = 0;  // one lane per row; wg entry slots per row
    int team = 1;         // waves sharing the walk (one workgroup)
    int *rk0 = nullptr;   // per lane: position of the row's first entry on this side in the factors
    int *rcnt = nullptr;  // per lane: its number of entries (0: no row)
    int *row = nullptr;   // row of the lane (-1: none)
    int *aux = nullptr;   // forward: the row's position in the backward stream; backward: its pivot's position
    int4 *meta = nullptr; // per lane: row, forward output position, ring slot of the result, the chunk's need
    unsigned *ent = nullptr;  // 16-byte entry records (column, value bits) packed after each factorisation
    float *dinv32 = nullptr;  // inverted pivots (backward), packed from the factors
    double *dinv64 = nullptr;
    double *pre64 = nullptr;  // forward: scale * r; backward: the forward result; in stream order
    float *pre32 = nullptr;
  } wave[2];  // [0] forward (L), [1] backward (U)
  bool wave_mode = false;
  int *wave_ecol = nullptr;  // per factor entry: the LDS ring slot of its dependency (level walks)
  // which sweep kernels apply the factors (gdp_solve_info.ilu_sweep)
  int sweep_kind() const { return wave_mode ? GDP_SWEEP_WAVE : GDP_SWEEP_FREE; }

  ~IluFactor() { release(); }
  void release();
  // Dependency levels and diagonal positions for pattern B (kept by pointer).
  int analyse(const DevCSR &B, const std::vector<int> &hia, const std::vector<int> &hja);
  // Take factors computed on the host, in their row layout; storage FP64
  // when f64, else FP32 (the values are converted once).
  template <typename T> int load(const IlutRows<T> &f, bool f64);
  // Factor values a (B's CSR layout); FP32 or FP64 storage.
  int factor(const double *a, bool f64, int *zero_pivot_row);
  // The same factorisation, enqueued only (no allocation, copy or host
  // synchronisation: usable under graph capture).  The storage must have
  // been allocated by reserve(); factor_status() reads the pivot flag.
  int reserve(bool f64);
  int factor_enqueue(const double *a);
  int factor_status(int *zero_pivot_row);
  // x = (LU)^{-1} (scale * r); may be enqueued repeatedly without host sync.
  // gate/gmode: skip when a Krylov control block says the work has ended.
  // addto (optional, distinct from x): also addto += x, in the same pass.
  template <typename TV>
  int apply(const TV *r, TV scale, TV *x, const int *gate = nullptr, int gmode = 0,
            TV *addto = nullptr);
  int reset_sweep_state();
  int check_timeout();
  unsigned long long wait_limit() const;
  template <typename T> void enqueue_levels(T *lu);

 private:
  int setup_sweeps(const int *hia, const int *hja);
  int setup_wave(const int *hia, const int *hja, const std::vector<int> &lrows,
                 const std::vector<int> &lptr);
  long wave_chunks_alloc(const WaveSweep &W) const;
  int wave_alloc_values();
  void wave_pack();
  void wave_release();
};

}  // namespace gdp
