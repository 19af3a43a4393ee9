// es_sieve_k3.hip -- sieve and batch kernel instances with 10..12 word masks (see
// es_sieve_dev.h).
#include "es_sieve_dev.h"

namespace es_dev {

KernelFn pick_kernel_p3(int F, uint32_t W, uint32_t G, uint32_t C) {
  switch (F) {
    case 10: return kernel_for<10>(W, G, C);
    case 11: return kernel_for<11>(W, G, C);
    case 12: return kernel_for<12>(W, G, C);
    default: return nullptr;
  }
}

GridBatchFn pick_grid_batch_p3(int F, uint32_t W, uint32_t G, uint32_t C) {
  switch (F) {
    case 10: return grid_batch_for<10>(W, G, C);
    case 11: return grid_batch_for<11>(W, G, C);
    case 12: return grid_batch_for<12>(W, G, C);
    default: return nullptr;
  }
}

RowsFn pick_rows_p3(int F, uint32_t C, uint32_t W, bool wave8) {
  switch (F) {
    case 10: return rows_kernel_for<10>(C, W, wave8);
    case 11: return rows_kernel_for<11>(C, W, wave8);
    default: return nullptr;
  }
}

RowsBatchFn pick_rows_batch_p3(int F, uint32_t C, uint32_t W, bool wave8) {
  switch (F) {
    case 10: return rows_batch_for<10>(C, W, wave8);
    case 11: return rows_batch_for<11>(C, W, wave8);
    default: return nullptr;
  }
}

}  // namespace es_dev
