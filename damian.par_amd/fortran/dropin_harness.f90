! dropin_harness.f90 -- compiles the drop-in replacement bodies of
! init_explicit / run_explicit (explicit_dmp.inc, the text INTEGRATION.md
! shows) against modules that declare the reference's globals with the
! reference's names and types (model.f90:22-97, io_utils.f90:5,
! communicator.f90:34, explicit.f90:8-12), and runs them on a site built the
! way the reference builds it on one rank.  The reference cannot be compiled
! here (it needs MPI, ParMETIS and m4-generated sources), so these modules
! stand in for the parts of it that the drop-in touches; save_step and
! save_mass write the reference's formats through the library's io.cpp-
! compatible writers.
!
!   dmp_dropin <outdir> <wavesv|pointd> [precision 0|1] [io_interval]
! writes <outdir>/step.000.bin (every io_interval steps, as save_step does)
! and <outdir>/mass.000.bin.
module model
    implicit none
    integer :: m_nummat, m_totnodes, m_totelems, m_nsteps, m_totrows, m_totcols
    real*8  :: m_duration, m_deltat, m_blength, m_mlength
    real*8, allocatable, dimension(:,:)   :: m_matconst
    integer :: m_nnodes, m_nelems, m_nrows, m_ncols
    integer, allocatable, dimension(:)    :: m_ematprop
    real*8, allocatable, dimension(:,:)   :: m_nodes
    integer, allocatable, dimension(:)    :: m_eptr
    integer, allocatable, dimension(:)    :: m_eind
    logical*1, allocatable, dimension(:)  :: m_ediv
    integer, allocatable, dimension(:,:)  :: m_hanging
    integer, allocatable, dimension(:)    :: m_absorbing
    character(len=6) :: m_prob
    real*8, dimension(2) :: m_worigin
    real*8               :: m_wttoorigin
    real*8               :: m_wangle
    real*8               :: m_wamp
    real*8               :: m_wfreq
    real*8, dimension(2) :: m_wvelo
    integer              :: m_wrows
    integer              :: m_wcols
    integer, allocatable, dimension(:,:) :: m_seismic
    real*8 :: m_pointamp
    real*8 :: m_pointsigma
    real*8 :: m_pointstart
    integer, allocatable, dimension(:) :: m_pointnodes
end module model

module communicator
    implicit none
    integer :: c_rank = 0
end module communicator

module utilities
    implicit none
contains
    subroutine error_exfem(msg)
        character(len=*), intent(in) :: msg
        write(*,*) msg
        error stop 1
    end subroutine
end module utilities

module io_utils
    use iso_c_binding
    use utilities
    use damian_step, only: dmp_write_step, dmp_write_mass
    implicit none
    integer :: io_interval
    logical*1 :: io_newsave = .TRUE.
    character(len=256) :: io_outdir
contains
    ! save_mass / save_step (io_utils.f90:334-405) with the file names of
    ! one rank
    subroutine save_mass(proc, mass)
        integer, intent(in)  :: proc
        real*8, dimension(:) :: mass
        character(len=256) :: filename
        integer(c_int) :: st
        write(filename, "(A,A6,I3.3,A4)") trim(io_outdir), "/mass.", proc, ".bin"
        st = dmp_write_mass(len_trim(filename), filename, size(mass, 1), mass)
        if (st /= 0) call error_exfem("save_mass failed")
    end subroutine

    subroutine save_step(proc, step, u, timer)
        integer, intent(in) :: proc, step
        real*8, intent(in), dimension(:) :: u
        integer, intent(inout) :: timer
        character(len=256) :: filename
        character(len=27)  :: elstr
        integer(c_int) :: st
        if ( mod(step, io_interval) == 0 ) then
            write(filename, "(A,A6,I3.3,A4)") trim(io_outdir), "/step.", proc, ".bin"
            if ( io_newsave ) then
                st = dmp_write_step(len_trim(filename), filename, "n", step/io_interval, size(u,1), u)
                io_newsave = .FALSE.
            else
                st = dmp_write_step(len_trim(filename), filename, "o", step/io_interval, size(u,1), u)
            end if
            if (st /= 0) call error_exfem("save_step failed")
            call end_timer(elstr, timer)
            call start_timer(timer)
        end if
    end subroutine

    subroutine start_timer(timer)
        integer, intent(out) :: timer
        call system_clock(timer)
    end subroutine

    subroutine end_timer(elstr, timer)
        character(len=27), intent(out) :: elstr
        integer, intent(in) :: timer
        integer :: now, rate
        call system_clock(now, rate)
        write(elstr, '(A,F12.3,A)') "Elapsed: ", real(now - timer, 8) / real(max(rate, 1), 8), " s"
    end subroutine
end module io_utils

module explicit
    use iso_c_binding
    use model
    use io_utils
    use communicator
    use utilities
    use damian_step, only: DMP_FP64
    implicit none
    real*8, dimension(:), allocatable  :: e_mfull
    real*8, dimension(:), allocatable  :: e_u
    real*8, dimension(:), allocatable  :: e_up
    type(c_ptr) :: e_dmp = c_null_ptr
    integer(c_int) :: dmp_precision = DMP_FP64
contains
    include 'explicit_dmp.inc'
end module explicit

program dmp_dropin
    use iso_c_binding
    use model
    use io_utils
    use explicit
    use f1_site
    implicit none
    character(len=256) :: arg, kind
    integer(c_int), allocatable :: eptr(:), eind(:), ematprop(:), absorbing(:), seismic(:,:)
    logical(c_bool), allocatable :: ediv(:)
    real(c_double), allocatable :: nodes(:,:), matconst(:,:)
    real(c_double) :: box(3), d1, d2
    integer(c_int) :: ns
    integer :: k, npt

    call get_command_argument(1, io_outdir)
    call get_command_argument(2, kind)
    if (command_argument_count() >= 3) then
        call get_command_argument(3, arg)
        read(arg, *) dmp_precision
    end if
    io_interval = 50
    if (command_argument_count() >= 4) then
        call get_command_argument(4, arg)
        read(arg, *) io_interval
    end if

    ! ---- what init_model leaves in the model module (one rank)
    call uniform_site(nodes, eptr, eind, ediv, ematprop, absorbing, seismic, ns, matconst, box)
    m_nnodes = size(nodes, 2)
    m_nelems = size(eptr) - 1
    m_nodes = nodes
    m_eptr = eptr
    m_eind = eind
    m_ediv = logical(ediv, kind=1)
    m_ematprop = ematprop
    m_nummat = 1
    m_matconst = matconst
    m_absorbing = absorbing
    allocate(m_hanging(3, 0))
    m_blength = h
    m_totcols = ncols
    m_totrows = nrows
    m_deltat = 0.002d0
    if (trim(kind) == "wavesv") then
        m_prob = "WAVESV"
        m_nsteps = 400
        m_seismic = seismic(:, 1:ns)
        m_worigin = (/ 100.0d0, 0.0d0 /)
        m_wttoorigin = 0.3d0
        m_wangle = 0.0d0
        m_wamp = 1.0d0
        m_wfreq = 4.0d0
        m_wvelo = (/ 1000.0d0, 500.0d0 /)
        m_wrows = ceiling(50.0d0 / h) + 1
        m_wcols = (ncols - (1 + ceiling(100.0d0 / h))) / 2
    else if (trim(kind) == "pointd") then
        ! setup_point_source (model.f90:1051-1078): nodes strictly closer than
        ! 6 m to (100, 0) or (60, 40), ascending
        m_prob = "POINTD"
        m_nsteps = 300
        m_pointamp = 1.0d-2
        m_pointsigma = 0.002d0
        m_pointstart = 0.15d0
        npt = 0
        allocate(m_pointnodes(m_nnodes))
        do k = 1, m_nnodes
            d1 = sqrt((100.0d0 - m_nodes(1, k))**2 + (0.0d0 - m_nodes(2, k))**2)
            d2 = sqrt((60.0d0 - m_nodes(1, k))**2 + (40.0d0 - m_nodes(2, k))**2)
            if (d1 < 6.0d0 .or. d2 < 6.0d0) then
                npt = npt + 1
                m_pointnodes(npt) = k
            end if
        end do
        m_pointnodes = m_pointnodes(1:npt)
    else
        call error_exfem("dmp_dropin: case must be wavesv or pointd")
    end if

    ! ---- solve() (main.f90:49-51)
    call init_explicit()
    call run_explicit()
end program dmp_dropin
