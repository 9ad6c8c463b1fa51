!-----------------------------------------------------------------------
! ISO_C_BINDING view of librsdft_gpu.so (include/rsdft_gpu.h) and
! drop-in wrappers that keep the reference routines' Fortran signatures:
!
!   gpu_hamiltonian(k,s,tpsi,htpsi,n1,n2,ib1,ib2)   hamiltonian_module.f90:20
!   gpu_op_kinetic(k,tpsi,htpsi,n1,n2,ib1,ib2)      kinetic_module.f90:186
!   gpu_op_localpot(s,mm,nn,f,vf)                   localpot_module.f90:27
!   gpu_op_nonlocal(k,s,tpsi,htpsi,n1,n2,ib1,ib2)   nonlocal_module.f90:20
!   gpu_conjugate_gradient(n1,n2,MB,k,s,unk,esp,res) cg_module.f90:54
!   gpu_gram_schmidt(unk,ML0,MB,n0,n1)              gram_schmidt_module.f90:28
!   gpu_subspace_diag(unk,esp,ML0,MB)               subspace_diag_module.f90:17
!
! A non-zero status is turned into `stop`, as the reference stops on errors
! (ps_nloc2_op_module.f90:155, subspace_solv_sl_module.f90:125-128).
! The op_* wrappers may be entered by every thread of an OpenMP parallel
! region (the reference calls them from inside one): the master thread
! drives the GPU and all threads meet at a barrier afterwards.
!-----------------------------------------------------------------------
module rsdft_gpu_capi
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: RSDFT_PREC_FP64 = 0
  integer(c_int), parameter :: RSDFT_PREC_MIXED_FP32 = 1
  integer(c_int), parameter :: RSDFT_PREC_MIXED_BF16 = 2

  interface
    integer(c_int) function rsdft_gpu_init(n1, n2, n3, hgrid, dv, md, coef_lap, coef_lap0) &
         bind(C, name="rsdft_gpu_init")
      import :: c_int, c_double
      integer(c_int), value :: n1, n2, n3, md
      real(c_double), intent(in) :: hgrid(3), coef_lap(*)
      real(c_double), value :: dv, coef_lap0
    end function
    integer(c_int) function rsdft_gpu_set_vloc(vloc, s) bind(C, name="rsdft_gpu_set_vloc")
      import :: c_int, c_double
      real(c_double), intent(in) :: vloc(*)
      integer(c_int), value :: s
    end function
    integer(c_int) function rsdft_gpu_set_projectors(nzlma, maxmjj, mjj, jjp, uvk, iuv) &
         bind(C, name="rsdft_gpu_set_projectors")
      import :: c_int, c_double
      integer(c_int), value :: nzlma, maxmjj
      integer(c_int), intent(in) :: mjj(*), jjp(*), iuv(*)
      real(c_double), intent(in) :: uvk(*)
    end function
    integer(c_int) function rsdft_gpu_set_projector_derivatives(natom, amap, mmjj, mjj_map, &
         jj_map, duvdr) bind(C, name="rsdft_gpu_set_projector_derivatives")
      import :: c_int, c_double
      integer(c_int), value :: natom, mmjj
      integer(c_int), intent(in) :: amap(*), mjj_map(*), jj_map(*)
      real(c_double), intent(in) :: duvdr(*)
    end function
    integer(c_int) function rsdft_gpu_set_precision(mode) bind(C, name="rsdft_gpu_set_precision")
      import :: c_int
      integer(c_int), value :: mode
    end function
    integer(c_int) function rsdft_gpu_set_solver_params(ncg, mloop) &
         bind(C, name="rsdft_gpu_set_solver_params")
      import :: c_int
      integer(c_int), value :: ncg, mloop
    end function
    integer(c_int) function rsdft_gpu_hamiltonian(k, s, tpsi, htpsi, n1, n2, ib1, ib2) &
         bind(C, name="rsdft_gpu_hamiltonian")
      import :: c_int, c_double
      integer(c_int), value :: k, s, n1, n2, ib1, ib2
      real(c_double), intent(in) :: tpsi(*)
      real(c_double), intent(inout) :: htpsi(*)
    end function
    integer(c_int) function rsdft_gpu_op_kinetic(k, tpsi, htpsi, n1, n2, ib1, ib2) &
         bind(C, name="rsdft_gpu_op_kinetic")
      import :: c_int, c_double
      integer(c_int), value :: k, n1, n2, ib1, ib2
      real(c_double), intent(in) :: tpsi(*)
      real(c_double), intent(inout) :: htpsi(*)
    end function
    integer(c_int) function rsdft_gpu_op_localpot(s, mm, nn, f, vf) &
         bind(C, name="rsdft_gpu_op_localpot")
      import :: c_int, c_double
      integer(c_int), value :: s, mm, nn
      real(c_double), intent(in) :: f(*)
      real(c_double), intent(inout) :: vf(*)
    end function
    integer(c_int) function rsdft_gpu_op_nonlocal(k, s, tpsi, htpsi, n1, n2, ib1, ib2) &
         bind(C, name="rsdft_gpu_op_nonlocal")
      import :: c_int, c_double
      integer(c_int), value :: k, s, n1, n2, ib1, ib2
      real(c_double), intent(in) :: tpsi(*)
      real(c_double), intent(inout) :: htpsi(*)
    end function
    integer(c_int) function rsdft_gpu_wf_attach(unk, ml0, mb) bind(C, name="rsdft_gpu_wf_attach")
      import :: c_int, c_double
      real(c_double), intent(inout) :: unk(*)
      integer(c_int), value :: ml0, mb
    end function
    integer(c_int) function rsdft_gpu_wf_sync(dir) bind(C, name="rsdft_gpu_wf_sync")
      import :: c_int
      integer(c_int), value :: dir
    end function
    integer(c_int) function rsdft_gpu_wf_detach() bind(C, name="rsdft_gpu_wf_detach")
      import :: c_int
    end function
    integer(c_int) function rsdft_gpu_conjugate_gradient(n1, n2, mb, k, s, unk, esp, res) &
         bind(C, name="rsdft_gpu_conjugate_gradient")
      import :: c_int, c_double
      integer(c_int), value :: n1, n2, mb, k, s
      real(c_double), intent(inout) :: unk(*), esp(*), res(*)
    end function
    integer(c_int) function rsdft_gpu_gram_schmidt(unk, ml0, mb, n0, n1) &
         bind(C, name="rsdft_gpu_gram_schmidt")
      import :: c_int, c_double
      real(c_double), intent(inout) :: unk(*)
      integer(c_int), value :: ml0, mb, n0, n1
    end function
    integer(c_int) function rsdft_gpu_subspace_diag(unk, esp, ml0, mb) &
         bind(C, name="rsdft_gpu_subspace_diag")
      import :: c_int, c_double
      real(c_double), intent(inout) :: unk(*), esp(*)
      integer(c_int), value :: ml0, mb
    end function
    integer(c_int) function rsdft_gpu_sweep(unk, esp, res, ml0, mb, nsweep) &
         bind(C, name="rsdft_gpu_sweep")
      import :: c_int, c_double
      real(c_double), intent(inout) :: unk(*), esp(*), res(*)
      integer(c_int), value :: ml0, mb, nsweep
    end function
    integer(c_int) function rsdft_gpu_calc_density(unk, ml0, mb, occ, rho) &
         bind(C, name="rsdft_gpu_calc_density")
      import :: c_int, c_double
      real(c_double), intent(in) :: unk(*), occ(*)
      real(c_double), intent(inout) :: rho(*)
      integer(c_int), value :: ml0, mb
    end function
    integer(c_int) function rsdft_gpu_force_ps_nloc2(unk, ml0, mb, occ, force2) &
         bind(C, name="rsdft_gpu_force_ps_nloc2")
      import :: c_int, c_double
      real(c_double), intent(in) :: unk(*), occ(*)
      real(c_double), intent(inout) :: force2(*)
      integer(c_int), value :: ml0, mb
    end function
    integer(c_int) function rsdft_gpu_band_energy_grad(unk, ml0, mb, occ, e, dedv, dedr) &
         bind(C, name="rsdft_gpu_band_energy_grad")
      import :: c_int, c_double
      real(c_double), intent(in) :: unk(*), occ(*)
      real(c_double), intent(inout) :: e, dedv(*), dedr(*)
      integer(c_int), value :: ml0, mb
    end function
    ! SCF iteration (scf_module.f90:146-366)
    integer(c_int) function rsdft_gpu_calc_hartree(rho, vh, e_h) bind(C, name="rsdft_gpu_calc_hartree")
      import :: c_int, c_double
      real(c_double), intent(in) :: rho(*)
      real(c_double), intent(out) :: vh(*), e_h
    end function
    integer(c_int) function rsdft_gpu_calc_xc(rho, vxc, e_xc) bind(C, name="rsdft_gpu_calc_xc")
      import :: c_int, c_double
      real(c_double), intent(in) :: rho(*)
      real(c_double), intent(out) :: vxc(*), e_xc(3)
    end function
    integer(c_int) function rsdft_gpu_calc_fermi(esp, nb, nelectron, ekbt, kinteg, occ, efermi) &
         bind(C, name="rsdft_gpu_calc_fermi")
      import :: c_int, c_double
      real(c_double), intent(in) :: esp(*)
      integer(c_int), value :: nb, kinteg
      real(c_double), value :: nelectron, ekbt
      real(c_double), intent(out) :: occ(*), efermi
    end function
    integer(c_int) function rsdft_gpu_scf_init(vion, rho, nelectron, mmix, beta, ekbt, kinteg) &
         bind(C, name="rsdft_gpu_scf_init")
      import :: c_int, c_double
      real(c_double), intent(in) :: vion(*), rho(*)
      real(c_double), value :: nelectron, beta, ekbt
      integer(c_int), value :: mmix, kinteg
    end function
    integer(c_int) function rsdft_gpu_scf_iterate(unk, ml0, mb, esp, occ, res, info) &
         bind(C, name="rsdft_gpu_scf_iterate")
      import :: c_int, c_double
      real(c_double), intent(inout) :: unk(*)
      integer(c_int), value :: ml0, mb
      real(c_double), intent(out) :: esp(*), occ(*), res(*), info(10)
    end function
    integer(c_int) function rsdft_gpu_scf_get_fields(rho, vloc, vh, vxc) &
         bind(C, name="rsdft_gpu_scf_get_fields")
      import :: c_int, c_double
      real(c_double), intent(out) :: rho(*), vloc(*), vh(*), vxc(*)
    end function
    ! allocation-free hamiltonian on device arrays (no reference counterpart)
    integer(c_int) function rsdft_gpu_hamiltonian_noalloc(k, s, d_tpsi, d_htpsi, n1, n2, ib1, ib2) &
         bind(C, name="rsdft_gpu_hamiltonian_noalloc")
      import :: c_int, c_ptr
      integer(c_int), value :: k, s, n1, n2, ib1, ib2
      type(c_ptr), value :: d_tpsi, d_htpsi
    end function
    ! complex wavefunctions at one k point: complex(8) arrays as the
    ! reference's complex unk(n1:n2, MB), zcoef_kin given as coef_nabk(3,Md)
    ! and const_k2, Bloch-phased uVk(maxmjj, nzlma)
    integer(c_int) function rsdft_gpu_set_kpoint(const_k2, coef_nabk, nzlma, maxmjj, mjj, jjp, &
         uvk, iuv) bind(C, name="rsdft_gpu_set_kpoint")
      import :: c_int, c_double, c_double_complex
      real(c_double), value :: const_k2
      real(c_double), intent(in) :: coef_nabk(*)
      integer(c_int), value :: nzlma, maxmjj
      integer(c_int), intent(in) :: mjj(*), jjp(*), iuv(*)
      complex(c_double_complex), intent(in) :: uvk(*)
    end function
    integer(c_int) function rsdft_gpu_zhamiltonian(k, s, tpsi, htpsi, n1, n2, ib1, ib2) &
         bind(C, name="rsdft_gpu_zhamiltonian")
      import :: c_int, c_double_complex
      integer(c_int), value :: k, s, n1, n2, ib1, ib2
      complex(c_double_complex), intent(in) :: tpsi(*)
      complex(c_double_complex), intent(out) :: htpsi(*)
    end function
    integer(c_int) function rsdft_gpu_zconjugate_gradient(unk, esp, res, ml0, mb) &
         bind(C, name="rsdft_gpu_zconjugate_gradient")
      import :: c_int, c_double, c_double_complex
      complex(c_double_complex), intent(inout) :: unk(*)
      real(c_double), intent(out) :: esp(*), res(*)
      integer(c_int), value :: ml0, mb
    end function
    integer(c_int) function rsdft_gpu_zgram_schmidt(unk, ml0, mb) &
         bind(C, name="rsdft_gpu_zgram_schmidt")
      import :: c_int, c_double_complex
      complex(c_double_complex), intent(inout) :: unk(*)
      integer(c_int), value :: ml0, mb
    end function
    integer(c_int) function rsdft_gpu_zsubspace_diag(unk, esp, ml0, mb) &
         bind(C, name="rsdft_gpu_zsubspace_diag")
      import :: c_int, c_double, c_double_complex
      complex(c_double_complex), intent(inout) :: unk(*)
      real(c_double), intent(out) :: esp(*)
      integer(c_int), value :: ml0, mb
    end function
    integer(c_int) function rsdft_gpu_zsweep(unk, esp, res, ml0, mb, nsweep) &
         bind(C, name="rsdft_gpu_zsweep")
      import :: c_int, c_double, c_double_complex
      complex(c_double_complex), intent(inout) :: unk(*)
      real(c_double), intent(out) :: esp(*), res(*)
      integer(c_int), value :: ml0, mb, nsweep
    end function
    integer(c_int) function rsdft_gpu_get_timers(t, n) bind(C, name="rsdft_gpu_get_timers")
      import :: c_int, c_double
      real(c_double), intent(inout) :: t(*)
      integer(c_int), value :: n
    end function
    integer(c_int) function rsdft_gpu_finalize() bind(C, name="rsdft_gpu_finalize")
      import :: c_int
    end function
    type(c_ptr) function rsdft_gpu_last_error() bind(C, name="rsdft_gpu_last_error")
      import :: c_ptr
    end function
  end interface

contains

  function rsdft_gpu_error_message() result(msg)
    character(len=1024) :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: i
    msg = ""
    call c_f_pointer(rsdft_gpu_last_error(), p, [1024])
    do i = 1, 1024
      if (p(i) == c_null_char) exit
      msg(i:i) = p(i)
    end do
  end function

  subroutine rsdft_gpu_check(ierr, name)
    integer(c_int), intent(in) :: ierr
    character(*), intent(in) :: name
    if (ierr /= 0) then
      write(*,*) "rsdft_gpu: ", name, " failed: ", trim(rsdft_gpu_error_message())
      stop "rsdft_gpu error"
    end if
  end subroutine

end module rsdft_gpu_capi


module rsdft_gpu_ops
  use rsdft_gpu_capi
  implicit none
  private
  public :: gpu_hamiltonian, gpu_op_kinetic, gpu_op_localpot, gpu_op_nonlocal
  public :: gpu_conjugate_gradient, gpu_gram_schmidt, gpu_subspace_diag

contains

  subroutine gpu_hamiltonian(k, s, tpsi, htpsi, n1, n2, ib1, ib2)
    integer, intent(IN) :: k, s, n1, n2, ib1, ib2
    real(8), intent(IN) :: tpsi(n1:n2, ib1:ib2)
    real(8), intent(OUT) :: htpsi(n1:n2, ib1:ib2)
    call rsdft_gpu_check(rsdft_gpu_hamiltonian(k, s, tpsi, htpsi, n1, n2, ib1, ib2), &
                         "hamiltonian")
  end subroutine

  subroutine gpu_op_kinetic(k, tpsi, htpsi, n1, n2, ib1, ib2)
    integer, intent(IN) :: k, n1, n2, ib1, ib2
    real(8), intent(IN) :: tpsi(n1:n2, ib1:ib2)
    real(8), intent(INOUT) :: htpsi(n1:n2, ib1:ib2)
    integer(c_int) :: ierr
    ierr = 0
!$OMP master
    ierr = rsdft_gpu_op_kinetic(k, tpsi, htpsi, n1, n2, ib1, ib2)
    call rsdft_gpu_check(ierr, "op_kinetic")
!$OMP end master
!$OMP barrier
  end subroutine

  subroutine gpu_op_localpot(s, mm, nn, f, vf)
    integer, intent(IN) :: s, mm, nn
    real(8), intent(IN) :: f(mm, nn)
    real(8), intent(INOUT) :: vf(mm, nn)
    integer(c_int) :: ierr
    ierr = 0
!$OMP master
    ierr = rsdft_gpu_op_localpot(s, mm, nn, f, vf)
    call rsdft_gpu_check(ierr, "op_localpot")
!$OMP end master
!$OMP barrier
  end subroutine

  subroutine gpu_op_nonlocal(k, s, tpsi, htpsi, n1, n2, ib1, ib2)
    integer, intent(IN) :: k, s, n1, n2, ib1, ib2
    real(8), intent(IN) :: tpsi(n1:n2, ib1:ib2)
    real(8), intent(INOUT) :: htpsi(n1:n2, ib1:ib2)
    integer(c_int) :: ierr
    ierr = 0
!$OMP master
    ierr = rsdft_gpu_op_nonlocal(k, s, tpsi, htpsi, n1, n2, ib1, ib2)
    call rsdft_gpu_check(ierr, "op_nonlocal")
!$OMP end master
!$OMP barrier
  end subroutine

  subroutine gpu_conjugate_gradient(n1, n2, MB, k, s, unk, esp, res)
    integer, intent(IN) :: n1, n2, MB, k, s
    real(8), intent(INOUT) :: unk(n1:n2, MB), esp(MB), res(MB)
    call rsdft_gpu_check(rsdft_gpu_conjugate_gradient(n1, n2, MB, k, s, unk, esp, res), &
                         "conjugate_gradient")
  end subroutine

  subroutine gpu_gram_schmidt(unk, ML0, MB, n0, n1)
    integer, intent(IN) :: ML0, MB, n0, n1
    real(8), intent(INOUT) :: unk(ML0, MB)
    call rsdft_gpu_check(rsdft_gpu_gram_schmidt(unk, ML0, MB, n0, n1), "gram_schmidt")
  end subroutine

  subroutine gpu_subspace_diag(unk, esp, ML0, MB)
    integer, intent(IN) :: ML0, MB
    real(8), intent(INOUT) :: unk(ML0, MB), esp(MB)
    call rsdft_gpu_check(rsdft_gpu_subspace_diag(unk, esp, ML0, MB), "subspace_diag")
  end subroutine

end module rsdft_gpu_ops
