! dmp_drive.f90 -- Fortran driver of the step library for a uniform site:
! 200 m x 100 m, 5 m quads, Vs 500, Vp 1000, rho 2000, vertical SV Ricker
! (f0 4 Hz, t0 0.3 s) through a 100 m x 50 m plane-wave box, 400 steps of
! 0.002 s, 49 surface receivers every 4.1 m.  Builds the model arrays the way
! the reference driver does on one rank (initial_mesh numbering,
! create_absorbingbc, wave_elems), then replaces init_explicit/run_explicit
! by library calls in output blocks.
!
!   dmp_drive <outdir> [precision 0|1]
! writes <outdir>/recv.bin (raw float64, 2 x nrecv per sample, samples at
! steps 0..400), <outdir>/step.000.bin and <outdir>/mass.000.bin in the
! reference's formats.
program dmp_drive
    use iso_c_binding
    use damian_step
    implicit none
    integer, parameter :: nrows = 21, ncols = 41, nsteps = 400, nblock = 100
    real(c_double), parameter :: h = 5.0d0, tol = 1.0d-9
    integer(c_int) :: nnodes, nelems, st, nsaved, prec, i, j, e, k, ns, nr, step
    integer(c_int), allocatable :: eptr(:), eind(:), ematprop(:), absorbing(:), seismic(:,:)
    integer(c_int), allocatable :: relem(:)
    integer(c_int) :: hang(3)
    logical(c_bool), allocatable :: ediv(:), rff(:)
    real(c_double), allocatable :: nodes(:,:), rxy(:,:), rec(:,:,:), blk(:), u(:), up(:), mf(:)
    real(c_double) :: matconst(5, 1), origin(2), box(3), left, right, top, x, dt, mu, lam
    integer :: wrows, wcols, nsamp
    type(c_ptr) :: hnd
    character(len=256) :: outdir, arg, fname
    integer(c_int) :: nf

    call get_command_argument(1, outdir)
    prec = 0
    if (command_argument_count() >= 2) then
        call get_command_argument(2, arg)
        read(arg, *) prec
    end if

    ! ---- model (setup_materials, setup_problem_params, init_wave)
    mu = 2000.0d0 * 500.0d0**2
    lam = 2000.0d0 * 1000.0d0**2 - 2.0d0 * mu
    matconst(:, 1) = (/ mu, lam, 2000.0d0, 1000.0d0, 500.0d0 /)
    wrows = ceiling(50.0d0 / h) + 1
    wcols = (ncols - (1 + ceiling(100.0d0 / h))) / 2
    left = real(wcols - 1, 8) * h
    right = real((ncols - wcols) - 1, 8) * h
    top = real(wrows - 1, 8) * h
    origin = (/ 100.0d0, 0.0d0 /)
    box = (/ left, right, top /)
    dt = 0.002d0

    ! ---- initial_mesh on one rank
    nnodes = nrows * ncols
    nelems = (nrows - 1) * (ncols - 1)
    allocate(nodes(2, nnodes), eptr(nelems + 1), eind(4 * nelems), ediv(4 * nelems))
    allocate(ematprop(nelems), absorbing(nnodes))
    k = 0
    do i = 1, ncols
        do j = 1, nrows
            k = k + 1
            nodes(:, k) = (/ real(i - 1, 8) * h, real(j - 1, 8) * h /)
        end do
    end do
    e = 0
    eptr(1) = 1
    do i = 1, ncols - 1
        do j = 1, nrows - 1
            e = e + 1
            eptr(e + 1) = eptr(e) + 4
            eind(eptr(e):eptr(e + 1) - 1) = (/ (i - 1) * nrows + j, i * nrows + j, &
                                               i * nrows + j + 1, (i - 1) * nrows + j + 1 /)
        end do
    end do
    ediv = .false.
    ematprop = 1

    ! ---- create_absorbingbc
    do k = 1, nnodes
        absorbing(k) = 0
        if (abs(nodes(1, k)) < tol .or. abs(nodes(1, k) - real(ncols - 1, 8) * h) < tol) then
            absorbing(k) = 2
        else if (abs(nodes(2, k) - real(nrows - 1, 8) * h) < tol) then
            absorbing(k) = 1
        end if
    end do

    ! ---- wave_elems (kinds SHORI=1, SVERL=2, SVERR=4, CORNR=5, CORNL=6)
    allocate(seismic(2, nelems))
    ns = 0
    do e = 1, nelems
        associate (n1 => nodes(:, eind(eptr(e))), n2 => nodes(:, eind(eptr(e) + 1)))
            k = 0
            if (abs(n2(1) - left) < tol .and. abs(n2(2) - top) < tol) then
                k = 6
            else if (abs(n1(1) - right) < tol .and. abs(n1(2) - top) < tol) then
                k = 5
            else if (abs(n2(1) - left) < tol .and. (n2(2) - top) < -tol) then
                k = 2
            else if (abs(n1(1) - right) < tol .and. (n1(2) - top) < -tol) then
                k = 4
            else if ((n2(1) - left) > tol .and. (n1(1) - right) < -tol .and. &
                     abs(n1(2) - top) < tol) then
                k = 1
            end if
        end associate
        if (k /= 0) then
            ns = ns + 1
            seismic(:, ns) = (/ e, k /)
        end if
    end do

    ! ---- surface receivers x = -3.3 + 4.1 i, i = 1..49, between the nodes
    !      (find_highest with SHEET -3.3 197 4.1)
    nr = int((197.0d0 + 3.3d0) / 4.1d0) + 1
    allocate(relem(nr), rxy(2, nr), rff(nr))
    do i = 1, nr
        x = -3.3d0 + 4.1d0 * real(i, 8)
        relem(i) = 0
        rxy(:, i) = (/ x, 0.0d0 /)
        do e = 1, nelems
            associate (n1 => nodes(:, eind(eptr(e))), n3 => nodes(:, eind(eptr(e) + 2)))
                if (abs(n1(2)) < tol .and. .not. ((x - n3(1)) > tol) .and. &
                    .not. ((x - n1(1)) < -tol)) relem(i) = e
            end associate
        end do
        rff(i) = ((x - left) < -tol) .or. ((x - right) > tol)
    end do

    ! ---- init_explicit + run_explicit through the library
    st = dmp_create(hnd, prec, 0_c_int)
    if (st /= DMP_OK) call fail(st)
    st = dmp_set_mesh(hnd, nelems, eptr, eind, ediv, nnodes, nodes, ematprop, 1_c_int, &
                      matconst, c_null_ptr)
    if (st /= DMP_OK) call fail(st)
    st = dmp_set_boundaries(hnd, absorbing, 0_c_int, hang, ns, seismic)
    if (st /= DMP_OK) call fail(st)
    st = dmp_set_wave(hnd, "WAVESV", origin, 0.3d0, 0.0d0, 1.0d0, 4.0d0, 1000.0d0, 500.0d0, box)
    if (st /= DMP_OK) call fail(st)
    st = dmp_set_time(hnd, dt, nsteps, 1_c_int)
    if (st /= DMP_OK) call fail(st)
    st = dmp_set_receivers(hnd, nr, relem, rxy, rff)
    if (st /= DMP_OK) call fail(st)
    st = dmp_init(hnd)
    if (st /= DMP_OK) call fail(st)

    allocate(rec(2, nr, nsteps + 1), blk(2 * nr * (nblock + 1)))
    nsamp = 0
    do step = 0, nsteps - 1, nblock
        st = dmp_run(hnd, nblock, blk, nsaved)
        if (st /= DMP_OK) call fail(st)
        rec(:, :, nsamp + 1:nsamp + nsaved) = reshape(blk(1:2 * nr * nsaved), (/ 2, nr, nsaved /))
        nsamp = nsamp + nsaved
    end do

    allocate(u(2 * nnodes), up(2 * nnodes), mf(2 * nnodes))
    st = dmp_get_state(hnd, u, up, mf)
    if (st /= DMP_OK) call fail(st)
    fname = trim(outdir) // "/step.000.bin"
    nf = len_trim(fname)
    st = dmp_write_step(nf, fname, "n", nsteps, 2 * nnodes, u)
    if (st /= DMP_OK) call fail(st)
    fname = trim(outdir) // "/mass.000.bin"
    nf = len_trim(fname)
    st = dmp_write_mass(nf, fname, 2 * nnodes, mf)
    if (st /= DMP_OK) call fail(st)
    open(unit=11, file=trim(outdir) // "/recv.bin", access="stream", form="unformatted", &
         status="replace")
    write(11) rec(:, :, 1:nsamp)
    close(11)
    st = dmp_destroy(hnd)
    write(*, '(A,I0,A,I0)') "dmp_drive: samples ", nsamp, " receivers ", nr

contains

    subroutine fail(code)
        integer(c_int), intent(in) :: code
        write(*, '(A,I0,A,A)') "dmp_drive: error ", code, ": ", trim(dmp_message(hnd))
        error stop 3
    end subroutine

end program dmp_drive
