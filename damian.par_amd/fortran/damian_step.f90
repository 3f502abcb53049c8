! damian_step.f90 -- ISO_C_BINDING interface to libdamian_hip.so
! (include/damian_step.h), one interface per exported entry point.
! Follows the reference's Fortran <-> C++ layer (io_utils.f90:14-286): every
! argument by reference, 1-based integer arrays passed unchanged, logical*1
! arrays as logical(c_bool), character data as c_char arrays with explicit
! lengths.  Every function returns an integer status (0 = success).
! Arguments the C side accepts as NULL are OPTIONAL here: an absent actual
! argument is passed as a null pointer (Fortran 2018, C interoperability of
! optional dummies).
module damian_step
    use iso_c_binding
    implicit none
    integer(c_int), parameter :: DMP_OK = 0, DMP_EINVAL = -1, DMP_ENOMEM = -2, &
                                 DMP_EHIP = -3, DMP_EUNSUPPORTED = -4, DMP_ESTATE = -5, &
                                 DMP_EWAVE = -6
    integer(c_int), parameter :: DMP_FP64 = 0, DMP_FP32C = 1

    interface
        ! ---- lifetime
        integer(c_int) function dmp_create(h, precision, device) bind(C, name="dmp_create")
            import :: c_int, c_ptr
            type(c_ptr), intent(out) :: h
            integer(c_int), intent(in) :: precision, device
        end function

        integer(c_int) function dmp_destroy(h) bind(C, name="dmp_destroy")
            import :: c_int, c_ptr
            type(c_ptr), intent(inout) :: h
        end function

        integer(c_int) function dmp_last_error(h, buf, buflen) bind(C, name="dmp_last_error")
            import :: c_int, c_ptr, c_char
            type(c_ptr), value :: h
            character(kind=c_char), dimension(*) :: buf
            integer(c_int), intent(in) :: buflen
        end function

        ! ---- model state (model.f90:34-97)
        integer(c_int) function dmp_set_mesh(h, nelems, eptr, eind, ediv, nnodes, nodes, &
                                             ematprop, nummat, matconst, elem_props) &
                                             bind(C, name="dmp_set_mesh")
            import :: c_int, c_ptr, c_double, c_bool
            type(c_ptr), value :: h
            integer(c_int), intent(in) :: nelems, nnodes, nummat
            integer(c_int), dimension(*), intent(in) :: eptr, eind
            logical(c_bool), dimension(*), intent(in) :: ediv
            real(c_double), dimension(*), intent(in) :: nodes
            integer(c_int), dimension(*), intent(in), optional :: ematprop
            real(c_double), dimension(*), intent(in), optional :: matconst, elem_props
        end function

        integer(c_int) function dmp_set_boundaries(h, absorbing, nhang, hanging, nseis, &
                                                   seismic) bind(C, name="dmp_set_boundaries")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
            integer(c_int), dimension(*), intent(in) :: absorbing
            integer(c_int), intent(in) :: nhang, nseis
            integer(c_int), dimension(*), intent(in), optional :: hanging, seismic
        end function

        integer(c_int) function dmp_set_wave(h, prob, origin, t0, angle, amp, freq, vp, vs, &
                                             box) bind(C, name="dmp_set_wave")
            import :: c_int, c_ptr, c_double, c_char
            type(c_ptr), value :: h
            character(kind=c_char), dimension(*), intent(in) :: prob
            real(c_double), dimension(*), intent(in) :: origin, box
            real(c_double), intent(in) :: t0, angle, amp, freq, vp, vs
        end function

        integer(c_int) function dmp_set_point_source(h, prob, amp, sigma, start, npoints, &
                                                     pointnodes) &
                                                     bind(C, name="dmp_set_point_source")
            import :: c_int, c_ptr, c_double, c_char
            type(c_ptr), value :: h
            character(kind=c_char), dimension(*), intent(in) :: prob
            real(c_double), intent(in) :: amp, sigma, start
            integer(c_int), intent(in) :: npoints
            integer(c_int), dimension(*), intent(in), optional :: pointnodes
        end function

        integer(c_int) function dmp_set_time(h, dt, nsteps, io_interval) &
                                             bind(C, name="dmp_set_time")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(in) :: dt
            integer(c_int), intent(in) :: nsteps, io_interval
        end function

        integer(c_int) function dmp_set_receivers(h, nrecv, recv_elem, recv_xy, recv_ff) &
                                                  bind(C, name="dmp_set_receivers")
            import :: c_int, c_ptr, c_double, c_bool
            type(c_ptr), value :: h
            integer(c_int), intent(in) :: nrecv
            integer(c_int), dimension(*), intent(in), optional :: recv_elem
            real(c_double), dimension(*), intent(in), optional :: recv_xy
            logical(c_bool), dimension(*), intent(in), optional :: recv_ff
        end function

        ! ---- the loop (explicit.f90:79-159)
        integer(c_int) function dmp_init(h) bind(C, name="dmp_init")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function

        integer(c_int) function dmp_run(h, nsteps_block, recv_out, nsaved) bind(C, name="dmp_run")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            integer(c_int), intent(in) :: nsteps_block
            real(c_double), dimension(*), intent(out), optional :: recv_out
            integer(c_int), intent(out), optional :: nsaved
        end function

        ! dmp_run without any allocation (library scratch of 16 MiB per GPU;
        ! see include/damian_step.h)
        integer(c_int) function dmp_run_noalloc(h, nsteps_block, recv_out, nsaved) &
                                                bind(C, name="dmp_run_noalloc")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            integer(c_int), intent(in) :: nsteps_block
            real(c_double), dimension(*), intent(out), optional :: recv_out
            integer(c_int), intent(out), optional :: nsaved
        end function

        integer(c_int) function dmp_get_state(h, u, up, mfull) bind(C, name="dmp_get_state")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), dimension(*), intent(out), optional :: u, up, mfull
        end function

        integer(c_int) function dmp_last_run_ms(h, ms) bind(C, name="dmp_last_run_ms")
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: h
            real(c_float), intent(out) :: ms
        end function

        integer(c_int) function dmp_batch_last_run_ms(b, ms) bind(C, name="dmp_batch_last_run_ms")
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: b
            real(c_float), intent(out) :: ms
        end function

        integer(c_int) function dmp_get_step(h, step) bind(C, name="dmp_get_step")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
            integer(c_int), intent(out) :: step
        end function

        integer(c_int) function dmp_export_mesh(h, eptr, eind, ediv, hanging, seismic) &
                                                bind(C, name="dmp_export_mesh")
            import :: c_int, c_ptr, c_bool
            type(c_ptr), value :: h
            integer(c_int), dimension(*), intent(out), optional :: eptr, eind, hanging, seismic
            logical(c_bool), dimension(*), intent(out), optional :: ediv
        end function

        ! ---- tuning and batching
        integer(c_int) function dmp_set_tuning(h, tuning) bind(C, name="dmp_set_tuning")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
            integer(c_int), dimension(4), intent(in), optional :: tuning
        end function

        integer(c_int) function dmp_get_tuning(h, tuning) bind(C, name="dmp_get_tuning")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
            integer(c_int), dimension(4), intent(out) :: tuning
        end function

        integer(c_int) function dmp_autotune(h, nsteps, tuning_out) bind(C, name="dmp_autotune")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
            integer(c_int), intent(in) :: nsteps
            integer(c_int), dimension(4), intent(out), optional :: tuning_out
        end function

        integer(c_int) function dmp_batch_create(b, n, handles) bind(C, name="dmp_batch_create")
            import :: c_int, c_ptr
            type(c_ptr), intent(out) :: b
            integer(c_int), intent(in) :: n
            type(c_ptr), dimension(*), intent(in) :: handles
        end function

        integer(c_int) function dmp_batch_destroy(b) bind(C, name="dmp_batch_destroy")
            import :: c_int, c_ptr
            type(c_ptr), intent(inout) :: b
        end function

        integer(c_int) function dmp_batch_run(b, nsteps, recv_out, nsaved) bind(C, name="dmp_batch_run")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: b
            integer(c_int), intent(in) :: nsteps
            real(c_double), dimension(*), intent(out), optional :: recv_out
            integer(c_int), intent(out), optional :: nsaved
        end function

        integer(c_int) function dmp_batch_get_state(b, instance, u, up) &
                                                    bind(C, name="dmp_batch_get_state")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: b
            integer(c_int), intent(in) :: instance
            real(c_double), dimension(*), intent(out), optional :: u, up
        end function

        integer(c_int) function dmp_batch_set_tuning(b, tuning) bind(C, name="dmp_batch_set_tuning")
            import :: c_int, c_ptr
            type(c_ptr), value :: b
            integer(c_int), dimension(4), intent(in), optional :: tuning
        end function

        integer(c_int) function dmp_batch_autotune(b, nsteps, tuning_out) &
                                                   bind(C, name="dmp_batch_autotune")
            import :: c_int, c_ptr
            type(c_ptr), value :: b
            integer(c_int), intent(in) :: nsteps
            integer(c_int), dimension(4), intent(out), optional :: tuning_out
        end function

        ! ---- output formats (io.cpp)
        integer(c_int) function dmp_write_step(nfname, fname, mode, step, ndofs, u) &
                                               bind(C, name="dmp_write_step")
            import :: c_int, c_double, c_char
            integer(c_int), intent(in) :: nfname, step, ndofs
            character(kind=c_char), dimension(*), intent(in) :: fname, mode
            real(c_double), dimension(*), intent(in) :: u
        end function

        integer(c_int) function dmp_write_mass(nfname, fname, ndofs, mass) &
                                               bind(C, name="dmp_write_mass")
            import :: c_int, c_double, c_char
            integer(c_int), intent(in) :: nfname, ndofs
            character(kind=c_char), dimension(*), intent(in) :: fname
            real(c_double), dimension(*), intent(in) :: mass
        end function

        integer(c_int) function dmp_write_mesh(nfname, fname, nelems, eptr, eind, ediv, nnodes, nodes) &
                                               bind(C, name="dmp_write_mesh")
            import :: c_int, c_double, c_char, c_bool
            integer(c_int), intent(in) :: nfname, nelems, nnodes
            character(kind=c_char), dimension(*), intent(in) :: fname
            integer(c_int), dimension(*), intent(in) :: eptr, eind
            logical(c_bool), dimension(*), intent(in) :: ediv
            real(c_double), dimension(*), intent(in) :: nodes
        end function

        integer(c_int) function dmp_write_tparams(nfname, fname, deltat, iointerval) &
                                                  bind(C, name="dmp_write_tparams")
            import :: c_int, c_double, c_char
            integer(c_int), intent(in) :: nfname, iointerval
            character(kind=c_char), dimension(*), intent(in) :: fname
            real(c_double), intent(in) :: deltat
        end function

        integer(c_int) function dmp_write_rank(nfname, fname, myrank) bind(C, name="dmp_write_rank")
            import :: c_int, c_char
            integer(c_int), intent(in) :: nfname, myrank
            character(kind=c_char), dimension(*), intent(in) :: fname
        end function
    end interface

contains

    ! Message of the last failure on h (a null handle: of the library), as a
    ! blank-padded Fortran string.
    function dmp_message(h) result(msg)
        type(c_ptr), intent(in) :: h
        character(len=256) :: msg
        character(kind=c_char), dimension(256) :: buf
        integer(c_int) :: st, i
        st = dmp_last_error(h, buf, 256_c_int)
        do i = 1, 256
            msg(i:i) = buf(i)
        end do
    end function

end module damian_step
