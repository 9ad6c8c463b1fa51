!-----------------------------------------------------------------------
! Fortran driver of the GPU library through rsdft_gpu_ops: it does what a
! calc_sweep caller does (sweep_module.f90:147-158) with the reference's
! routine signatures.  Reads one stream-unformatted input file (grid, FD
! coefficients, Vloc, projector tables, psi), applies the hamiltonian once,
! orthonormalises, runs nsweep CG -> GS -> subspace sweeps, and writes
! H psi(initial), esp, res, the final psi and H psi(initial) once more, from
! the op_* wrappers entered by four OpenMP threads, to the output file.
!   usage: sweep_driver <input> <output>
!-----------------------------------------------------------------------
program sweep_driver
  use rsdft_gpu_capi
  use rsdft_gpu_ops
  implicit none
  integer :: N(3), Md, nzlma, maxmjj, MB, nsweep, ML, it
  real(8) :: h(3), dV, coef_lap0
  real(8), allocatable :: coef_lap(:,:), vloc(:), uVk(:,:), unk(:,:), hunk(:,:), hunk2(:,:)
  real(8), allocatable :: esp(:), res(:)
  integer, allocatable :: MJJ(:), JJP(:,:), iuV(:)
  character(len=1024) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) N, Md, nzlma, maxmjj, MB, nsweep
  read(10) h, dV, coef_lap0
  ML = N(1) * N(2) * N(3)
  allocate(coef_lap(3, Md), vloc(ML), MJJ(nzlma), JJP(maxmjj, nzlma), uVk(maxmjj, nzlma))
  allocate(iuV(nzlma), unk(ML, MB), hunk(ML, MB), esp(MB), res(MB))
  read(10) coef_lap, vloc, MJJ, JJP, uVk, iuV, unk
  close(10)

  call rsdft_gpu_check(rsdft_gpu_init(N(1), N(2), N(3), h, dV, Md, coef_lap, coef_lap0), "init")
  call rsdft_gpu_check(rsdft_gpu_set_vloc(vloc, 1), "set_vloc")
  call rsdft_gpu_check(rsdft_gpu_set_projectors(nzlma, maxmjj, MJJ, JJP, uVk, iuV), &
                       "set_projectors")

  call gpu_hamiltonian(1, 1, unk, hunk, 1, ML, 1, MB)
  ! the same operator the way hamiltonian_module.f90:38-70 calls it: every
  ! thread of an OpenMP region enters op_kinetic, op_localpot, op_nonlocal
  allocate(hunk2(ML, MB))
  hunk2 = 0.d0
!$OMP parallel num_threads(4) default(shared)
  call gpu_op_kinetic(1, unk, hunk2, 1, ML, 1, MB)
  call gpu_op_localpot(1, ML, MB, unk, hunk2)
  call gpu_op_nonlocal(1, 1, unk, hunk2, 1, ML, 1, MB)
!$OMP end parallel
  call gpu_gram_schmidt(unk, ML, MB, 1, MB)
  esp = 0.d0
  res = 0.d0
  do it = 1, nsweep
    call gpu_conjugate_gradient(1, ML, MB, 1, 1, unk, esp, res)
    call gpu_gram_schmidt(unk, ML, MB, 1, MB)
    call gpu_subspace_diag(unk, esp, ML, MB)
  end do

  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) hunk, esp, res, unk, hunk2
  close(11)
  call rsdft_gpu_check(rsdft_gpu_finalize(), "finalize")
end program sweep_driver
