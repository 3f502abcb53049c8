! dmp_drive.f90 -- Fortran driver of the step library on the uniform site of
! f1_site (200 m x 100 m, 5 m quads, 400 steps of 0.002 s, 49 surface
! receivers): builds the model arrays the way the reference driver does on
! one rank, then replaces init_explicit/run_explicit by library calls in
! output blocks of 100 steps.
!
!   dmp_drive <outdir> [precision 0|1] [single|batch]
! single (default): vertical SV Ricker (f0 4 Hz, t0 0.3 s); writes
!   <outdir>/recv.bin (raw float64, 2 x nrecv per sample, steps 0..400),
!   <outdir>/step.000.bin and <outdir>/mass.000.bin in the reference's formats.
! batch: the same SV instance and a P wave at 30 deg on the same site, set up
!   but not initialised one by one, advanced together by one batch; writes
!   <outdir>/recv_batch.bin (per sample: instance 0's receivers, then
!   instance 1's) and <outdir>/step.00<q>.bin with instance q's final u.
program dmp_drive
    use iso_c_binding
    use damian_step
    use f1_site
    implicit none
    integer, parameter :: nsteps = 400, nblock = 100
    integer(c_int) :: nnodes, nelems, st, nsaved, prec, ns, nr, step, q, nf, ninst
    integer(c_int), allocatable :: eptr(:), eind(:), ematprop(:), absorbing(:), seismic(:,:)
    integer(c_int), allocatable :: relem(:)
    logical(c_bool), allocatable :: ediv(:), rff(:)
    real(c_double), allocatable :: nodes(:,:), matconst(:,:), rxy(:,:), rec(:,:,:), blk(:)
    real(c_double), allocatable :: u(:), up(:), mf(:)
    real(c_double) :: origin(2), box(3)
    integer :: nsamp
    type(c_ptr) :: hnd(2), bat
    character(len=256) :: outdir, arg, fname, mode
    character(len=6) :: prob(2)
    real(c_double) :: angle(2)

    call get_command_argument(1, outdir)
    prec = 0
    if (command_argument_count() >= 2) then
        call get_command_argument(2, arg)
        read(arg, *) prec
    end if
    mode = "single"
    if (command_argument_count() >= 3) call get_command_argument(3, mode)

    call uniform_site(nodes, eptr, eind, ediv, ematprop, absorbing, seismic, ns, matconst, box)
    call sheet_receivers(nodes, eptr, eind, box, nr, relem, rxy, rff)
    nnodes = size(nodes, 2)
    nelems = size(eptr) - 1
    origin = (/ 100.0d0, 0.0d0 /)
    prob = (/ "WAVESV", "WAVEP " /)
    angle = (/ 0.0d0, 30.0d0 * (acos(-1.0d0) / 180.0d0) /)
    ninst = 1
    if (trim(mode) == "batch") ninst = 2

    ! ---- init_explicit: model arrays to the library (set up, not yet initialised)
    do q = 1, ninst
        st = dmp_create(hnd(q), prec, 0_c_int)
        call check(st, hnd(q))
        st = dmp_set_mesh(hnd(q), nelems, eptr, eind, ediv, nnodes, nodes, ematprop, 1_c_int, &
                          matconst)
        call check(st, hnd(q))
        st = dmp_set_boundaries(hnd(q), absorbing, 0_c_int, nseis=ns, seismic=seismic)
        call check(st, hnd(q))
        st = dmp_set_wave(hnd(q), prob(q), origin, 0.3d0, angle(q), 1.0d0, 4.0d0, 1000.0d0, &
                          500.0d0, box)
        call check(st, hnd(q))
        st = dmp_set_time(hnd(q), 0.002d0, nsteps, 1_c_int)
        call check(st, hnd(q))
        st = dmp_set_receivers(hnd(q), nr, relem, rxy, rff)
        call check(st, hnd(q))
    end do

    if (ninst == 1) then
        st = dmp_init(hnd(1))
        call check(st, hnd(1))
        allocate(rec(2, nr, nsteps + 1), blk(2 * nr * (nblock + 1)))
        nsamp = 0
        do step = 0, nsteps - 1, nblock
            st = dmp_run(hnd(1), nblock, blk, nsaved)
            call check(st, hnd(1))
            rec(:, :, nsamp + 1:nsamp + nsaved) = reshape(blk(1:2 * nr * nsaved), (/ 2, nr, nsaved /))
            nsamp = nsamp + nsaved
        end do
        allocate(u(2 * nnodes), up(2 * nnodes), mf(2 * nnodes))
        st = dmp_get_state(hnd(1), u, up, mf)
        call check(st, hnd(1))
        fname = trim(outdir) // "/step.000.bin"
        nf = len_trim(fname)
        st = dmp_write_step(nf, fname, "n", nsteps, 2 * nnodes, u)
        call check(st, c_null_ptr)
        fname = trim(outdir) // "/mass.000.bin"
        nf = len_trim(fname)
        st = dmp_write_mass(nf, fname, 2 * nnodes, mf)
        call check(st, c_null_ptr)
        open(unit=11, file=trim(outdir) // "/recv.bin", access="stream", form="unformatted", &
             status="replace")
        write(11) rec(:, :, 1:nsamp)
        close(11)
        st = dmp_destroy(hnd(1))
        write(*, '(A,I0,A,I0)') "dmp_drive: samples ", nsamp, " receivers ", nr
    else
        ! ---- both instances in one batch: one launch per step block for both
        st = dmp_batch_create(bat, ninst, hnd)
        call check(st, c_null_ptr)
        allocate(rec(2, ninst * nr, nsteps + 1), blk(2 * ninst * nr * (nblock + 1)))
        nsamp = 0
        do step = 0, nsteps - 1, nblock
            st = dmp_batch_run(bat, nblock, blk, nsaved)
            call check(st, c_null_ptr)
            rec(:, :, nsamp + 1:nsamp + nsaved) = &
                reshape(blk(1:2 * ninst * nr * nsaved), (/ 2, ninst * nr, nsaved /))
            nsamp = nsamp + nsaved
        end do
        allocate(u(2 * nnodes))
        do q = 1, ninst
            st = dmp_batch_get_state(bat, q - 1, u=u)
            call check(st, c_null_ptr)
            write(fname, '(A,A,I3.3,A)') trim(outdir), "/step.", q - 1, ".bin"
            nf = len_trim(fname)
            st = dmp_write_step(nf, fname, "n", nsteps, 2 * nnodes, u)
            call check(st, c_null_ptr)
        end do
        open(unit=11, file=trim(outdir) // "/recv_batch.bin", access="stream", &
             form="unformatted", status="replace")
        write(11) rec(:, :, 1:nsamp)
        close(11)
        st = dmp_batch_destroy(bat)
        do q = 1, ninst
            st = dmp_destroy(hnd(q))
        end do
        write(*, '(A,I0,A,I0)') "dmp_drive: batch samples ", nsamp, " receivers ", ninst * nr
    end if

contains

    ! A failed call: its message (a null handle gives the library's last one).
    subroutine check(code, h)
        integer(c_int), intent(in) :: code
        type(c_ptr), intent(in) :: h
        if (code == DMP_OK) return
        write(*, '(A,I0,A,A)') "dmp_drive: error ", code, ": ", trim(dmp_message(h))
        error stop 3
    end subroutine

end program dmp_drive
