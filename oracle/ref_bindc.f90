! ORACLE (test infrastructure only): bind(C) entry points around the
! reference's own Fortran kernels, compiled from the reference sources into
! oracle/_ref/ by oracle/Makefile.  Used only to pin oracle/fortran_kernels.c.
! Buffers are Fortran-ordered (column-major); the Python side passes
! np.asfortranarray copies, exactly like the f2py wrappers.

subroutine ref_diff_2d(c, dc, n, m) bind(C, name="ref_diff_2d")
    use iso_c_binding
    integer(c_int), value :: n, m
    real(c_double), intent(in) :: c(n, m)
    real(c_double), intent(out) :: dc(n, m)
    interface
        subroutine diff_2d(c, dc, n, m)
            integer, intent(in) :: n
            integer :: m
            real(8), intent(in) :: c(n, m)
            real(8), intent(out) :: dc(n, m)
        end subroutine
    end interface
    dc = 0.0d0
    call diff_2d(c, dc, n, m)
end subroutine

subroutine ref_tdma_2d(a, b, c, d, k, x, n, m) bind(C, name="ref_tdma_2d")
    use iso_c_binding
    integer(c_int), value :: k, n, m
    real(c_double), intent(in) :: a(n - k), b(n), c(n - k), d(n, m)
    real(c_double), intent(out) :: x(n, m)
    interface
        subroutine solve_tdma_2d(a, b, c, d, k, x, n, m)
            integer, intent(in) :: n
            real(8), intent(in) :: a(:), b(:), c(:)
            integer, intent(in) :: k
            integer :: m
            real(8), intent(in) :: d(n, m)
            real(8), intent(out) :: x(n, m)
        end subroutine
    end interface
    x = 0.0d0
    call solve_tdma_2d(a, b, c, d, k, x, n, m)
end subroutine

subroutine ref_fdma_2d(l, d, u1, u2, x, axis, n, m) bind(C, name="ref_fdma_2d")
    use iso_c_binding
    integer(c_int), value :: axis, n, m
    real(c_double), intent(in) :: l(*), d(*), u1(*), u2(*)
    real(c_double), intent(inout) :: x(n, m)
    integer :: nn
    interface
        subroutine solve_fdma_2d(l, d, u1, u2, x, axis, n, m)
            integer, intent(in) :: axis, n, m
            real(8), intent(in) :: d(:), u1(:), u2(:), l(:)
            real(8), intent(inout) :: x(n, m)
        end subroutine
    end interface
    if (axis == 0) then
        nn = n
    else
        nn = m
    end if
    call solve_fdma_2d(l(1:nn - 2), d(1:nn), u1(1:nn - 2), u2(1:nn - 4), x, axis, n, m)
end subroutine

subroutine ref_init_fdma(A, d, u1, u2, l, n) bind(C, name="ref_init_fdma")
    use iso_c_binding
    integer(c_int), value :: n
    real(c_double), intent(in) :: A(n, n)
    real(c_double), intent(out) :: d(n), u1(n - 2), u2(n - 4), l(n - 2)
    interface
        subroutine init_fdma(A, d, u1, u2, l, n)
            integer, intent(in) :: n
            real(8), intent(in) :: A(n, n)
            real(8), intent(out) :: d(n), u1(n - 2), u2(n - 4), l(n - 2)
        end subroutine
    end interface
    call init_fdma(A, d, u1, u2, l, n)
end subroutine

subroutine ref_fdma_type2(A, C, lam, x, axis, singular, n, m) bind(C, name="ref_fdma_type2")
    use iso_c_binding
    integer(c_int), value :: axis, singular, n, m
    real(c_double), intent(in) :: A(*), C(*), lam(*)
    real(c_double), intent(inout) :: x(n, m)
    integer :: nn, nl
    logical :: sg
    interface
        subroutine solve_fdma_type2(A, C, lam, x, axis, singular, n, m)
            integer, intent(in) :: axis, n, m
            real(8), intent(in) :: A(:, :), C(:, :), lam(:)
            real(8), intent(inout) :: x(n, m)
            logical, intent(in) :: singular
        end subroutine
    end interface
    if (axis == 0) then
        nn = n
        nl = m
    else
        nn = m
        nl = n
    end if
    sg = (singular /= 0)
    call solve_fdma_type2(reshape(A(1:nn * nn), [nn, nn]), reshape(C(1:nn * nn), [nn, nn]), &
                          lam(1:nl), x, axis, sg, n, m)
end subroutine

subroutine ref_twodma_2d(d, u1, x, axis, n, m) bind(C, name="ref_twodma_2d")
    use iso_c_binding
    integer(c_int), value :: axis, n, m
    real(c_double), intent(in) :: d(*), u1(*)
    real(c_double), intent(inout) :: x(n, m)
    integer :: nn
    interface
        subroutine solve_twodma_2d(d, u1, x, axis, n, m)
            integer, intent(in) :: n, m, axis
            real(8), intent(in) :: d(:), u1(:)
            real(8), intent(inout) :: x(n, m)
        end subroutine
    end interface
    if (axis == 0) then
        nn = n
    else
        nn = m
    end if
    call solve_twodma_2d(d(1:nn), u1(1:nn - 2), x, axis, n, m)
end subroutine
