// es_sieve_k0.hip -- sieve and batch kernel instances with 0..3 word masks (see
// es_sieve_dev.h).
#include "es_sieve_dev.h"

namespace es_dev {

KernelFn pick_kernel_p0(int F, uint32_t W, uint32_t G, uint32_t C) {
  switch (F) {
    case 0: return kernel_for<0>(W, G, C);
    case 1: return kernel_for<1>(W, G, C);
    case 2: return kernel_for<2>(W, G, C);
    case 3: return kernel_for<3>(W, G, C);
    default: return nullptr;
  }
}

GridBatchFn pick_grid_batch_p0(int F, uint32_t W, uint32_t G, uint32_t C) {
  switch (F) {
    case 0: return grid_batch_for<0>(W, G, C);
    case 1: return grid_batch_for<1>(W, G, C);
    case 2: return grid_batch_for<2>(W, G, C);
    case 3: return grid_batch_for<3>(W, G, C);
    default: return nullptr;
  }
}

RowsFn pick_rows_p0(int F, uint32_t C, uint32_t W, bool wave8) {
  switch (F) {
    case 1: return rows_kernel_for<1>(C, W, wave8);
    case 2: return rows_kernel_for<2>(C, W, wave8);
    case 3: return rows_kernel_for<3>(C, W, wave8);
    default: return nullptr;
  }
}

RowsBatchFn pick_rows_batch_p0(int F, uint32_t C, uint32_t W, bool wave8) {
  switch (F) {
    case 1: return rows_batch_for<1>(C, W, wave8);
    case 2: return rows_batch_for<2>(C, W, wave8);
    case 3: return rows_batch_for<3>(C, W, wave8);
    default: return nullptr;
  }
}

}  // namespace es_dev
