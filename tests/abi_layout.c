/* Prints sizeof/offsetof of the C-ABI structs (compared with the ctypes
 * declarations by tests/test_cpu_abi.py). */
#include <stddef.h>
#include <stdio.h>
#include "../include/hipacc_hip.h"
#define F(T, m) printf("\"%s.%s\": %zu,\n", #T, #m, offsetof(T, m))
int main(void) {
  printf("{\n");
  F(hipacc_view, data); F(hipacc_view, dtype); F(hipacc_view, stride); F(hipacc_view, offset_y);
  F(hipacc_view, batch); F(hipacc_view, plane_stride);
  F(hipacc_local_desc, out); F(hipacc_local_desc, mask_w); F(hipacc_local_desc, weights_f32);
  F(hipacc_local_desc, domain); F(hipacc_local_desc, const_val); F(hipacc_local_desc, fp_contract);
  F(hipacc_separable_desc, row_f32); F(hipacc_separable_desc, acc_dtype);
  F(hipacc_separable_desc, const_val); F(hipacc_separable_desc, fp_contract);
  F(hipacc_point_desc, out); F(hipacc_point_desc, op); F(hipacc_point_desc, cvt);
  F(hipacc_bilateral_desc, ksize); F(hipacc_bilateral_desc, spatial);
  F(hipacc_bilateral_desc, const_val); F(hipacc_bilateral_desc, cvt);
  F(hipacc_local_grad_desc, grad_out); F(hipacc_local_grad_desc, grad_in);
  F(hipacc_local_grad_desc, mask_w); F(hipacc_local_grad_desc, weights_dev);
  F(hipacc_local_grad_desc, grad_weights_dev); F(hipacc_local_grad_desc, const_val);
  F(hipacc_local_grad_desc, workspace);
  F(hipacc_bilateral_grad_desc, norm); F(hipacc_bilateral_grad_desc, grad_in);
  F(hipacc_bilateral_grad_desc, ksize); F(hipacc_bilateral_grad_desc, spatial);
  F(hipacc_bilateral_grad_desc, const_val); F(hipacc_bilateral_grad_desc, grad_sigma_r_dev);
  F(hipacc_bilateral_grad_desc, workspace);
  F(hipacc_tune_entry, weights_dev); F(hipacc_tune_entry, size_class);
  F(hipacc_tune_entry, variant); F(hipacc_tune_entry, ms);
  printf("\"sizeof.tune\": %zu,\n", sizeof(hipacc_tune_entry));
  printf("\"sizeof.lgrad\": %zu, \"sizeof.bgrad\": %zu,\n", sizeof(hipacc_local_grad_desc),
         sizeof(hipacc_bilateral_grad_desc));
  printf("\"sizeof.view\": %zu, \"sizeof.local\": %zu, \"sizeof.sep\": %zu, \"sizeof.point\": %zu, \"sizeof.bilat\": %zu\n}\n",
         sizeof(hipacc_view), sizeof(hipacc_local_desc), sizeof(hipacc_separable_desc),
         sizeof(hipacc_point_desc), sizeof(hipacc_bilateral_desc));
  return 0;
}
