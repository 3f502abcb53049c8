! f1_site.f90 -- the uniform test site of the Fortran drivers, built the way
! the reference does on one rank: 200 m x 100 m of 5 m quads (initial_mesh
! numbering, mesh_utils.f90:80-110: node (col-1)*rows + row, corners BL, BR,
! TR, TL), Vs 500, Vp 1000, rho 2000, Lysmer flags of create_absorbingbc
! (model.f90:1080-1096), the 100 m x 50 m plane-wave box of init_wave
! (model.f90:982-1001) and its layer of wave_elems (model.f90:1015-1045).
module f1_site
    use iso_c_binding
    implicit none
    integer, parameter :: nrows = 21, ncols = 41
    real(c_double), parameter :: h = 5.0d0, tol = 1.0d-9

contains

    subroutine uniform_site(nodes, eptr, eind, ediv, ematprop, absorbing, seismic, ns, &
                            matconst, box)
        real(c_double), allocatable, intent(out) :: nodes(:,:), matconst(:,:)
        integer(c_int), allocatable, intent(out) :: eptr(:), eind(:), ematprop(:), absorbing(:), &
                                                    seismic(:,:)
        logical(c_bool), allocatable, intent(out) :: ediv(:)
        integer(c_int), intent(out) :: ns
        real(c_double), intent(out) :: box(3)
        integer :: nnodes, nelems, i, j, k, e, wrows, wcols
        real(c_double) :: mu, lam, left, right, top

        mu = 2000.0d0 * 500.0d0**2
        lam = 2000.0d0 * 1000.0d0**2 - 2.0d0 * mu
        allocate(matconst(5, 1))
        matconst(:, 1) = (/ mu, lam, 2000.0d0, 1000.0d0, 500.0d0 /)
        wrows = ceiling(50.0d0 / h) + 1
        wcols = (ncols - (1 + ceiling(100.0d0 / h))) / 2
        left = real(wcols - 1, 8) * h
        right = real((ncols - wcols) - 1, 8) * h
        top = real(wrows - 1, 8) * h
        box = (/ left, right, top /)

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

        do k = 1, nnodes
            absorbing(k) = 0
            if (abs(nodes(1, k)) < tol .or. abs(nodes(1, k) - real(ncols - 1, 8) * h) < tol) then
                absorbing(k) = 2
            else if (abs(nodes(2, k) - real(nrows - 1, 8) * h) < tol) then
                absorbing(k) = 1
            end if
        end do

        ! wave_elems (kinds SHORI=1, SVERL=2, SVERR=4, CORNR=5, CORNL=6)
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
    end subroutine

    ! Surface receivers x = -3.3 + 4.1 i, i = 1..49, between the nodes
    ! (find_highest with SHEET -3.3 197 4.1, visualization.f90:286-305), free
    ! field added outside the box (inside_soil, model.f90:217-235).
    subroutine sheet_receivers(nodes, eptr, eind, box, nr, relem, rxy, rff)
        real(c_double), intent(in) :: nodes(:,:), box(3)
        integer(c_int), intent(in) :: eptr(:), eind(:)
        integer(c_int), intent(out) :: nr
        integer(c_int), allocatable, intent(out) :: relem(:)
        real(c_double), allocatable, intent(out) :: rxy(:,:)
        logical(c_bool), allocatable, intent(out) :: rff(:)
        integer :: i, e
        real(c_double) :: x
        nr = int((197.0d0 + 3.3d0) / 4.1d0) + 1
        allocate(relem(nr), rxy(2, nr), rff(nr))
        do i = 1, nr
            x = -3.3d0 + 4.1d0 * real(i, 8)
            relem(i) = 0
            rxy(:, i) = (/ x, 0.0d0 /)
            do e = 1, size(eptr) - 1
                associate (n1 => nodes(:, eind(eptr(e))), n3 => nodes(:, eind(eptr(e) + 2)))
                    if (abs(n1(2)) < tol .and. .not. ((x - n3(1)) > tol) .and. &
                        .not. ((x - n1(1)) < -tol)) relem(i) = e
                end associate
            end do
            rff(i) = ((x - box(1)) < -tol) .or. ((x - box(2)) > tol)
        end do
    end subroutine

end module f1_site
