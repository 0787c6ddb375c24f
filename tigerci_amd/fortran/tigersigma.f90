! Fortran 2003 binding of the tigersigma C ABI (include/tigersigma.h).
!
! A Fortran Davidson driver (davidson_liu, davidson.f90:47) reaches the GPU
! sigma build through these interfaces; the drop-in htimesc body that uses
! them is shown in the integration notes of this repository.  All sizes are
! explicit 64-bit integers, so the module compiles unchanged with or without
! -fdefault-integer-8 (the reference builds with 8-byte default integers).
module tigersigma
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: TCI_OK = 0, TCI_ERR_ARG = 1, TCI_ERR_HIP = 2, &
                                       TCI_ERR_OOM = 3, TCI_ERR_NUMERIC = 4
  integer(c_int32_t), parameter, public :: TCI_F64 = 0, TCI_F32 = 1
  integer(c_int32_t), parameter, public :: TCI_SIGMA_FP64 = 0, TCI_SIGMA_BF16X3 = 1, TCI_SIGMA_FP16X3 = 2
  ! acpf_flag values of the reference input (gtimesc.f90 get_gvalues)
  integer(c_int32_t), parameter, public :: TCI_ACPF = 1, TCI_ACPF2 = 2, TCI_MRAQCC = 3, TCI_ACPF_CUSTOM = 4, &
                                           TCI_ACPF_TEST = 9
  integer(c_int32_t), parameter, public :: TCI_ACCUMULATE = 256, TCI_ADD_EREF = 512, TCI_SIGMA_GRAPH = 1024
  integer(c_int32_t), parameter, public :: TCI_DAVIDSON_GRAPH = 2048

  type, bind(c), public :: tci_problem
    integer(c_int64_t) :: n_int, n_v, n_chol
    type(c_ptr) :: eps            ! n_int + n_v orbital energies
    type(c_ptr) :: chol           ! (n_orb(n_orb+1)/2) records of n_chol values
    integer(c_int32_t) :: chol_dtype, ladder_dtype
    integer(c_int64_t) :: n_pairs
    type(c_ptr) :: pairs          ! n_pairs x 2 int32, (i, j) with i >= j, 0-based
    type(c_ptr) :: dom_off        ! n_pairs + 1 int64
    type(c_ptr) :: dom_idx        ! int32 sorted virtual indices, 0-based
    type(c_ptr) :: h              ! c_null_ptr, or the packed one-electron integrals pp(n_orb(n_orb+1)/2)
  end type tci_problem

  type, bind(c), public :: tci_plan_info
    integer(c_int64_t) :: n_csf, n_ladder_items
    real(c_double) :: e_ref, flops_ladder, flops_sigma, bytes_ladder
    integer(c_int64_t) :: device_bytes
  end type tci_plan_info

  type, bind(c), public :: tci_davidson_params
    integer(c_int32_t) :: roots, maxit
    real(c_double) :: energy_tol, norm_tol
    integer(c_int32_t) :: precision
    real(c_double) :: switch_tol
    integer(c_int32_t) :: max_vectors, verbose
  end type tci_davidson_params

  type, bind(c), public :: tci_davidson_result
    integer(c_int32_t) :: iterations, sigma_fp64, sigma_mixed, converged
    real(c_double) :: seconds, seconds_sigma
    integer(c_int32_t) :: collapses
  end type tci_davidson_result

  public :: tci_plan_create, tci_plan_destroy, tci_plan_get_info, tci_sigma_apply, tci_sigma_apply_batch, &
            tci_htimesc, tci_diagonal, tci_davidson_run, tci_davidson_gen_run, tci_acpf_gvalues, &
            tci_device_alloc, tci_device_free, tci_memcpy_h2d, tci_memcpy_d2h, tci_memcpy_d2d, &
            tci_device_synchronize, tci_profile_enable, tci_profile_read, tci_plan_tuning, tci_selftest_mfma, &
            tci_last_error_string

  interface
    integer(c_int) function tci_plan_create(prob, plan) bind(c, name="tci_plan_create")
      import :: c_int, c_ptr, tci_problem
      type(tci_problem), intent(in) :: prob
      type(c_ptr), intent(out) :: plan
    end function
    subroutine tci_plan_destroy(plan) bind(c, name="tci_plan_destroy")
      import :: c_ptr
      type(c_ptr), value :: plan
    end subroutine
    integer(c_int) function tci_plan_get_info(plan, info) bind(c, name="tci_plan_get_info")
      import :: c_int, c_ptr, tci_plan_info
      type(c_ptr), value :: plan
      type(tci_plan_info), intent(out) :: info
    end function
    ! device vectors c, sigma; flags = tci_precision + TCI_ACCUMULATE / TCI_ADD_EREF / TCI_SIGMA_GRAPH
    integer(c_int) function tci_sigma_apply(plan, c, sigma, n, flags, stream) bind(c, name="tci_sigma_apply")
      import :: c_int, c_ptr, c_int64_t, c_int32_t
      type(c_ptr), value :: plan, c, sigma, stream
      integer(c_int64_t), value :: n
      integer(c_int32_t), value :: flags
    end function
    ! nvec device vectors back to back in c and sigma (not overlapping)
    integer(c_int) function tci_sigma_apply_batch(plan, c, sigma, n, nvec, flags, stream) &
        bind(c, name="tci_sigma_apply_batch")
      import :: c_int, c_ptr, c_int64_t, c_int32_t
      type(c_ptr), value :: plan, c, sigma, stream
      integer(c_int64_t), value :: n
      integer(c_int32_t), value :: nvec, flags
    end function
    ! host vectors, htimesc semantics: sigmavec += H civec
    integer(c_int) function tci_htimesc(plan, civec, sigmavec, n) bind(c, name="tci_htimesc")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: plan
      real(c_double), intent(in) :: civec(*)
      real(c_double), intent(inout) :: sigmavec(*)
      integer(c_int64_t), value :: n
    end function
    integer(c_int) function tci_diagonal(plan, diag, n, stream) bind(c, name="tci_diagonal")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: plan, diag, stream
      integer(c_int64_t), value :: n
    end function
    integer(c_int) function tci_davidson_run(plan, params, energies, vectors, result) &
        bind(c, name="tci_davidson_run")
      import :: c_int, c_ptr, c_double, tci_davidson_params, tci_davidson_result
      type(c_ptr), value :: plan, vectors
      type(tci_davidson_params), intent(in) :: params
      real(c_double), intent(out) :: energies(*)
      type(tci_davidson_result), intent(out) :: result
    end function

    ! davidson_liu_gen_eig: (H - E_ref) c = E G c, g_values(5) from tci_acpf_gvalues
    integer(c_int) function tci_davidson_gen_run(plan, params, g_values, energies, vectors, result) &
        bind(c, name="tci_davidson_gen_run")
      import :: c_int, c_ptr, c_double, tci_davidson_params, tci_davidson_result
      type(c_ptr), value :: plan, vectors
      type(tci_davidson_params), intent(in) :: params
      real(c_double), intent(in) :: g_values(5)
      real(c_double), intent(out) :: energies(*)
      type(tci_davidson_result), intent(out) :: result
    end function

    integer(c_int) function tci_acpf_gvalues(functional, n_corr_elec, custom_g, g_values) &
        bind(c, name="tci_acpf_gvalues")
      import :: c_int, c_int32_t, c_double
      integer(c_int32_t), value :: functional, n_corr_elec
      real(c_double), value :: custom_g
      real(c_double), intent(out) :: g_values(5)
    end function
    integer(c_int) function tci_device_alloc(bytes, ptr) bind(c, name="tci_device_alloc")
      import :: c_int, c_ptr, c_int64_t
      integer(c_int64_t), value :: bytes
      type(c_ptr), intent(out) :: ptr
    end function
    integer(c_int) function tci_device_free(ptr) bind(c, name="tci_device_free")
      import :: c_int, c_ptr
      type(c_ptr), value :: ptr
    end function
    integer(c_int) function tci_memcpy_h2d(dst, src, bytes) bind(c, name="tci_memcpy_h2d")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: dst, src
      integer(c_int64_t), value :: bytes
    end function
    integer(c_int) function tci_memcpy_d2h(dst, src, bytes) bind(c, name="tci_memcpy_d2h")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: dst, src
      integer(c_int64_t), value :: bytes
    end function
    integer(c_int) function tci_memcpy_d2d(dst, src, bytes) bind(c, name="tci_memcpy_d2d")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: dst, src
      integer(c_int64_t), value :: bytes
    end function
    integer(c_int) function tci_device_synchronize() bind(c, name="tci_device_synchronize")
      import :: c_int
    end function
    ! HIP-event timing of every sigma build: ladder and whole-build milliseconds per precision
    integer(c_int) function tci_profile_enable(plan, on) bind(c, name="tci_profile_enable")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: plan
      integer(c_int32_t), value :: on
    end function
    integer(c_int) function tci_profile_read(plan, precision, ladder_ms, ladder_launches, sigma_ms, sigma_calls) &
        bind(c, name="tci_profile_read")
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: plan
      integer(c_int32_t), value :: precision
      real(c_double), intent(out) :: ladder_ms, sigma_ms
      integer(c_int64_t), intent(out) :: ladder_launches, sigma_calls
    end function
    ! autotuning record (text, NUL-terminated, at most len-1 characters)
    integer(c_int) function tci_plan_tuning(plan, buf, len) bind(c, name="tci_plan_tuning")
      import :: c_int, c_ptr, c_char, c_int64_t
      type(c_ptr), value :: plan
      character(kind=c_char), intent(out) :: buf(*)
      integer(c_int64_t), value :: len
    end function
    integer(c_int) function tci_selftest_mfma() bind(c, name="tci_selftest_mfma")
      import :: c_int
    end function
    ! internal walk table (DRT walks in lexical order, keep flags); capacity 0 queries the count
    integer(c_int) function tci_internal_walks(n_int, n_v, n_elec, spin_mult, n_ref, references, n_frozen, n_ext, &
        capacity, occ, kept, n_walks) bind(c, name="tci_internal_walks")
      import :: c_int, c_int32_t, c_int64_t, c_int8_t
      integer(c_int32_t), value :: n_int, n_v, n_elec, spin_mult, n_ref, n_frozen, n_ext
      integer(c_int8_t), intent(in) :: references(*)
      integer(c_int64_t), value :: capacity
      integer(c_int8_t), intent(out) :: occ(*), kept(*)
      integer(c_int64_t), intent(out) :: n_walks
    end function
    ! allowed-virtual lists of walks under WP/TOV (find_allowed_virtuals)
    integer(c_int) function tci_allowed_virtuals(n_int, n_v, n_ref, references, n_walks, walk_occ, wp, tov, status, &
        list_of_walk, list_off, list_virt, list_capacity, n_lists) bind(c, name="tci_allowed_virtuals")
      import :: c_int, c_int32_t, c_int64_t, c_int8_t
      integer(c_int32_t), value :: n_int, n_v, n_ref
      integer(c_int8_t), intent(in) :: references(*), walk_occ(*), wp(*), tov(*)
      integer(c_int64_t), value :: n_walks, list_capacity
      integer(c_int32_t), intent(out) :: status(*), list_virt(*)
      integer(c_int64_t), intent(out) :: list_of_walk(*), list_off(*), n_lists
    end function
    ! AO -> MO transform of Cholesky vectors (records in the tci_problem chol layout)
    integer(c_int) function tci_cholesky_ao_to_mo(n_ao, n_mo, n_chol, n_ao_pairs, ao_pairs, ao_vec, mo_coef, &
        mo_records, norms) bind(c, name="tci_cholesky_ao_to_mo")
      import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
      integer(c_int32_t), value :: n_ao, n_mo
      integer(c_int64_t), value :: n_chol, n_ao_pairs
      integer(c_int32_t), intent(in) :: ao_pairs(*)
      real(c_double), intent(in) :: ao_vec(*), mo_coef(*)
      real(c_double), intent(out) :: mo_records(*)
      type(c_ptr), value :: norms
    end function
    integer(c_int) function tci_spin_tables(n_open, spin_mult, sft, spin_matrix, cycles, transpositions) &
        bind(c, name="tci_spin_tables")
      import :: c_int, c_int32_t, c_ptr
      integer(c_int32_t), value :: n_open, spin_mult
      type(c_ptr), value :: sft, spin_matrix, cycles, transpositions
    end function
    integer(c_int) function tci_contracted_cycles(n_open, spin_mult, n_rec, records, out_off, out) &
        bind(c, name="tci_contracted_cycles")
      import :: c_int, c_int32_t, c_int64_t, c_double
      integer(c_int32_t), value :: n_open, spin_mult
      integer(c_int64_t), value :: n_rec
      integer(c_int32_t), intent(in) :: records(*)
      integer(c_int64_t), intent(in) :: out_off(*)
      real(c_double), intent(out) :: out(*)
    end function
    integer(c_int) function tci_ladder_records(n_int, n_v, spin_mult, n_walks, head_weight, num_singles, list_off, &
        list_virt, ignorable, capacity, records, n_records) bind(c, name="tci_ladder_records")
      import :: c_int, c_int32_t, c_int64_t, c_ptr
      integer(c_int32_t), value :: n_int, n_v, spin_mult
      integer(c_int64_t), value :: n_walks, capacity
      integer(c_int64_t), intent(in) :: head_weight(*), list_off(*)
      integer(c_int32_t), intent(in) :: num_singles(*), list_virt(*)
      type(c_ptr), value :: ignorable, records
      integer(c_int64_t), intent(out) :: n_records
    end function
    type(c_ptr) function tci_last_error_c() bind(c, name="tci_last_error")
      import :: c_ptr
    end function
  end interface

contains

  ! text of the last error as a Fortran string (trailing blanks trimmed by the caller)
  function tci_last_error_string() result(msg)
    character(len=512) :: msg
    character(kind=c_char), pointer :: p(:)
    type(c_ptr) :: cp
    integer :: k
    msg = ' '
    cp = tci_last_error_c()
    if (.not. c_associated(cp)) return
    call c_f_pointer(cp, p, [512])
    do k = 1, 512
      if (p(k) == c_null_char) exit
      msg(k:k) = p(k)
    end do
  end function tci_last_error_string

end module tigersigma
