/*
 * bus.c -- shared CPU-bus helpers of the memory sidecar, CPU restatement.
 * TEST INFRASTRUCTURE ONLY (checker for nightstream_amd; see neo_oracle.h).
 *
 * Restates neo-memory/src/ajtai.rs:31-55 (decode_vector: z[c] = sum_rho b^rho Z[rho][c]),
 * neo-memory/src/cpu/bus_layout.rs:63-80 (bus_cell = bus_base + col_id * chunk_size + j,
 * time_index = m_in + j), neo-fold/src/memory_sidecar/cpu_bus.rs:969-1016
 * (decode_cpu_z_to_k, build_time_sparse_from_bus_col) and cpu_bus.rs:183-320
 * (chi_for_row_index, append_bus_openings_to_me_instance: y[col][rho] =
 * sum_j chi_r(m_in + j) Z[rho][bus_cell(col, j)], y padded to 64, y_scalar = sum_rho b^rho y[rho]).
 *
 * Pinned by: tests/test_oracle_bus.py -- decode against Python big integers; each time column
 * against a direct gather of the decoded witness; each opening row against the multilinear
 * extension (kmath) of that digit row's time vector at r, and y_scalar against the MLE of the
 * decoded time column built by build_time_sparse_from_bus_col.
 */
#include <stdlib.h>
#include "neo_oracle.h"

void or_bus_decode(const int8_t* Z, size_t m, fq* z) {
  for (size_t c = 0; c < m; c++) {
    fq acc = 0, pw = 1;
    for (size_t rho = 0; rho < NEO_D; rho++) {
      int8_t d = Z[rho * m + c];
      if (d > 0) acc = or_fq_add(acc, pw);
      else if (d < 0) acc = or_fq_sub(acc, pw);
      pw = or_fq_add(pw, pw);
    }
    z[c] = acc;
  }
}

long or_bus_time_col(const fq* z, size_t zlen, size_t m_in, size_t bus_base, size_t chunk, size_t bus_cols,
                     size_t col_id, size_t steps_len, size_t pow2_cycle, uint64_t* idx_out, kx* vals_out) {
  if (col_id >= bus_cols || steps_len > chunk) return -1;
  long n = 0;
  for (size_t j = 0; j < chunk; j++) {
    size_t t = m_in + j;
    if (t >= pow2_cycle) return -1;
    if (j >= steps_len) continue;
    size_t idx = bus_base + col_id * chunk + j;
    if (idx >= zlen) return -1;
    if (z[idx] != 0) {
      idx_out[n] = t;
      vals_out[n].c0 = z[idx];
      vals_out[n].c1 = 0;
      n++;
    }
  }
  return n;
}

int or_bus_openings(const int8_t* Z, size_t m, size_t m_in, size_t bus_base, size_t chunk, size_t bus_cols,
                    const kx* r, size_t ell_n, kx* y, kx* y_scalars) {
  size_t n_pad = (size_t)1 << ell_n;
  if (ell_n == 0) return -1;
  for (size_t j = 0; j < chunk; j++)
    if (m_in + j >= n_pad) return -1;
  if (bus_base + bus_cols * chunk > m) return -1;
  kx* w = (kx*)malloc((chunk + 1) * sizeof(kx));
  for (size_t j = 0; j < chunk; j++) {
    size_t idx = m_in + j;
    kx acc = {1, 0};
    for (size_t b = 0; b < ell_n; b++) {
      kx f = ((idx >> b) & 1) ? r[b] : or_k_sub((kx){1, 0}, r[b]);
      acc = or_k_mul(acc, f);
    }
    w[j] = acc;
  }
  for (size_t col = 0; col < bus_cols; col++) {
    kx* row = y + col * 64;
    for (size_t rho = 0; rho < 64; rho++) row[rho] = (kx){0, 0};
    for (size_t rho = 0; rho < NEO_D; rho++) {
      kx acc = {0, 0};
      for (size_t j = 0; j < chunk; j++) {
        int8_t d = Z[rho * m + bus_base + col * chunk + j];
        if (d > 0) acc = or_k_add(acc, w[j]);
        else if (d < 0) acc = or_k_sub(acc, w[j]);
      }
      row[rho] = acc;
    }
    kx s = {0, 0};
    fq pw = 1;
    for (size_t rho = 0; rho < NEO_D; rho++) {
      s = or_k_add(s, or_k_scale(row[rho], pw));
      pw = or_fq_add(pw, pw);
    }
    y_scalars[col] = s;
  }
  free(w);
  return 0;
}
