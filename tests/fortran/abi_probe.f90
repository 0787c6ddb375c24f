! Links the Fortran binding against libtigersigma and exercises calls that need
! no GPU: argument validation of tci_plan_create, the last-error text and the
! ACPF g values (host arithmetic).
program abi_probe
  use, intrinsic :: iso_c_binding
  use tigersigma
  implicit none
  type(tci_problem) :: prob
  type(tci_plan_info) :: info
  type(tci_davidson_params) :: dparams
  type(tci_davidson_result) :: dres
  type(c_ptr) :: plan
  integer(c_int) :: rc
  real(c_double), target :: eps(3)
  real(c_double) :: g(5)
  eps = [-1.0d0, 1.0d0, 2.0d0]
  prob%n_int = 0_c_int64_t      ! no internal orbitals: rejected before any device work
  prob%n_v = 2_c_int64_t
  prob%n_chol = 1_c_int64_t
  prob%eps = c_loc(eps)
  prob%chol = c_null_ptr
  prob%chol_dtype = TCI_F64
  prob%ladder_dtype = TCI_F64
  prob%n_pairs = 0_c_int64_t
  prob%pairs = c_null_ptr
  prob%dom_off = c_null_ptr
  prob%dom_idx = c_null_ptr
  prob%h = c_null_ptr
  rc = tci_plan_create(prob, plan)
  write (*, '(a,i0)') 'status ', rc
  write (*, '(a,a)') 'error ', trim(tci_last_error_string())
  write (*, '(a,4(1x,i0))') 'sizes', c_sizeof(prob), c_sizeof(info), c_sizeof(dparams), c_sizeof(dres)
  rc = tci_acpf_gvalues(TCI_ACPF2, 14_c_int32_t, 0.0d0, g)
  write (*, '(a,i0,5(1x,f8.5))') 'acpf2 ', rc, g
end program abi_probe
