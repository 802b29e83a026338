"""CPU oracle for the Fast-Higashi decomposition path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (the `fasthigashi` package, the
HIP library, `bench.py` outside its `cpu_baseline` leg) may import, call or link
anything under this directory.  It is the checker: a numpy restatement of the
reference algorithm, parameterised by dtype (float32 follows the reference op for
op; float64 is the FP64 oracle).

Pinning status of each function (what independent check fixes it):

* sort_indices / slice_rows / tile_plan / tile_coo / densify: the reference's own
  `Sparse` known-answer test (sparse_for_schic.py:634-661), the notebook's
  compact-width known answers (PFC tutorial.ipynb:1560-1581), scipy coo_tocsr and
  dense brute force (tests/test_oracle_packing.py).
* pack_chrom: brute force dense restatement on tiny inputs + invariants
  (FastHigashi_Wrapper.py:299,353-354) (tests/test_oracle_packing.py).
* partial_rwr: closed form Q_n = (rP)^n + (1-r) sum_{t<n} (rP)^t and its limit
  (1-r)(I-rP)^{-1}, identity at zero epochs, zero-column diagonal fill
  (tests/test_oracle_rwr.py).
* project2orthogonal: uniqueness of the polar factor for full column rank,
  orthonormality, symmetric PSD of Y^T polar(Y), S.sum() = trace((Y^T Y)^{1/2})
  (tests/test_oracle_linalg.py).
* parafac (CP-ALS): exact low-rank fixed point, normal equations after each
  factor update, balance_norm invariants (tests/test_oracle_linalg.py).
* sweep / fit / transform: every contraction checked against a float64
  brute-force einsum, the error decomposition checked against the directly
  evaluated PARAFAC2 loss (tests/test_oracle_sweep.py).
* parafac2_loss / parafac2_loss_grad: central finite differences of the loss for every
  factor entry of small random cases (non-orthonormal P too), zero loss and zero
  gradient at an exactly representable model, and for orthonormal P and M the
  reference's norm decomposition ||X||^2 - 2 <P^T X M, [[A, B, D]]> + core norm
  (parafac2_intergrative.py:621-632,667-668) (tests/test_oracle_loss.py).
* init_params: parity unpinned for the random draws (the reference never seeds,
  parafac2_intergrative.py:71-80,257-258); the oracle and the product share an
  injected init instead.
"""
