! fdriver.f90 -- test entry points (BIND(C), loaded with ctypes by the GPU
! tests) that call the Fortran mirrors of hierarquicos_amd/fortran the way the
! reference's Fortran drivers call octree_mod, api_mod, direct_forces and the
! Verlet / Parareal routines.  Each entry point returns the first library
! status since fd_reset (0 = success); errors do not stop the process.
MODULE fdriver
    USE ISO_C_BINDING
    USE hq_iface, ONLY: hq_stop_on_error, hq_status, HQ_OK
    IMPLICIT NONE
CONTAINS

    SUBROUTINE fd_reset() BIND(C, name="fd_reset")
        hq_stop_on_error = .FALSE.
        hq_status = HQ_OK
    END SUBROUTINE

    INTEGER(C_INT) FUNCTION fd_status() BIND(C, name="fd_status")
        fd_status = hq_status
    END FUNCTION

    ! Sizes and field offsets of the BIND(C) structs, for comparison with the
    ! C layout: out = [sizeof(params), offset(fine_theta2), offset(coarse_scheme),
    ! sizeof(info), offset(ms_build), offset(max_stack)].
    SUBROUTINE fd_layout(out) BIND(C, name="fd_layout")
        USE hq_iface, ONLY: hq_parareal_params, hq_tree_info_t
        INTEGER(C_INT64_T), INTENT(OUT) :: out(6)
        TYPE(hq_parareal_params), TARGET :: prm
        TYPE(hq_tree_info_t), TARGET :: inf
        INTEGER(C_INTPTR_T) :: b
        b = TRANSFER(C_LOC(prm), b)
        out(1) = C_SIZEOF(prm)
        out(2) = TRANSFER(C_LOC(prm % fine_theta2), b) - b
        out(3) = TRANSFER(C_LOC(prm % coarse_scheme), b) - b
        b = TRANSFER(C_LOC(inf), b)
        out(4) = C_SIZEOF(inf)
        out(5) = TRANSFER(C_LOC(inf % ms_build), b) - b
        out(6) = TRANSFER(C_LOC(inf % max_stack), b) - b
    END SUBROUTINE

    ! api3d.f90 surface: F(N,3) = compute_forces(..., quad), Fd(N,3) = compute_forces_direct.
    INTEGER(C_INT) FUNCTION fd_api(n, m, x, y, z, theta, G, eps, quad, F, Fd) BIND(C, name="fd_api")
        USE hq_api_mod, ONLY: compute_forces, api_direct => compute_forces_direct
        INTEGER(C_INT64_T), VALUE :: n
        REAL(C_DOUBLE), INTENT(IN) :: m(n), x(n), y(n), z(n)
        REAL(C_DOUBLE), VALUE :: theta, G, eps
        INTEGER(C_INT), VALUE :: quad
        REAL(C_DOUBLE), INTENT(OUT) :: F(n, 3), Fd(n, 3)
        CALL fd_reset()
        F = compute_forces(m, x, y, z, theta, G, eps, quad /= 0)
        Fd = api_direct(m, x, y, z, G, eps)
        fd_api = hq_status
    END FUNCTION

    ! An OctreeType of `method` as test_forces.f90:180-221 drives it: per-particle
    ! forces (methods 10-12, over nt OpenMP threads), dehnen_eval (20-22) or
    ! forces_groups (30-32).  F is (3,N).
    INTEGER(C_INT) FUNCTION fd_octree(n, m, x, y, z, method, gs, theta2, G, eps2, nt, F) &
            BIND(C, name="fd_octree")
        USE hq_octree_mod, ONLY: OctreeType
        INTEGER(C_INT64_T), VALUE :: n
        REAL(C_DOUBLE), INTENT(IN) :: m(n), x(n), y(n), z(n)
        INTEGER(C_INT), VALUE :: method, gs, nt
        REAL(C_DOUBLE), VALUE :: theta2, G, eps2
        REAL(C_DOUBLE), INTENT(OUT) :: F(3, n)
        CLASS(OctreeType), ALLOCATABLE :: tree
        INTEGER :: p, nn

        CALL fd_reset()
        F = 0.0_C_DOUBLE
        ALLOCATE(tree)
        CALL tree % pre_init(m, INT(method), group_size=INT(gs))
        CALL tree % init(x, y, z)
        nn = INT(n)
        SELECT CASE (method)
            CASE (10, 11, 12)
                IF (nt <= 1) THEN
                    DO p = 1, nn
                        F(:, p) = tree % forces(p, theta2, G, eps2)
                    END DO
                ELSE
                    !$OMP PARALLEL DO SHARED(F, tree) PRIVATE(p) NUM_THREADS(nt) SCHEDULE(DYNAMIC)
                    DO p = 1, nn
                        F(:, p) = tree % forces(p, theta2, G, eps2)
                    END DO
                    !$OMP END PARALLEL DO
                END IF
            CASE (20, 21, 22)
                CALL tree % dehnen_eval(theta2, eps2, G, F)
            CASE (30, 31, 32)
                CALL tree % forces_groups(theta2, G, eps2, F, 1)
        END SELECT
        DEALLOCATE(tree)
        fd_octree = hq_status
    END FUNCTION

    ! The bunched walk and the FMM of one tree (method 10 build), both (3,N).
    INTEGER(C_INT) FUNCTION fd_bunched_fmm(n, m, x, y, z, theta2, G, eps2, precision, order, &
                                           fmm_mode, Fb, Ff) BIND(C, name="fd_bunched_fmm")
        USE hq_octree_mod, ONLY: OctreeType
        INTEGER(C_INT64_T), VALUE :: n
        REAL(C_DOUBLE), INTENT(IN) :: m(n), x(n), y(n), z(n)
        REAL(C_DOUBLE), VALUE :: theta2, G, eps2
        INTEGER(C_INT), VALUE :: precision, order, fmm_mode
        REAL(C_DOUBLE), INTENT(OUT) :: Fb(3, n), Ff(3, n)
        TYPE(OctreeType) :: tree
        CALL fd_reset()
        CALL tree % pre_init(m, 10)
        CALL tree % init(x, y, z)
        CALL tree % forces_bunched(theta2, G, eps2, Fb, INT(precision))
        CALL tree % fmm(INT(order), theta2, G, eps2, Ff, INT(fmm_mode))
        CALL tree % destroy()
        fd_bunched_fmm = hq_status
    END FUNCTION

    ! direct_forces.f90: F(3,N) = compute_forces_direct(m, x, y, z, G, eps2, nt).
    INTEGER(C_INT) FUNCTION fd_direct(n, m, x, y, z, G, eps2, F) BIND(C, name="fd_direct")
        USE hq_direct_mod, ONLY: compute_forces_direct
        INTEGER(C_INT64_T), VALUE :: n
        REAL(C_DOUBLE), INTENT(IN) :: m(n), x(n), y(n), z(n)
        REAL(C_DOUBLE), VALUE :: G, eps2
        REAL(C_DOUBLE), INTENT(OUT) :: F(3, n)
        CALL fd_reset()
        F = compute_forces_direct(m, x, y, z, G, eps2, 1)
        fd_direct = hq_status
    END FUNCTION

    ! test_forces.f90:66-127 on (3,N) qs, F, Fd.
    INTEGER(C_INT) FUNCTION fd_errors(n, m, qs, F, Fd, errors) BIND(C, name="fd_errors")
        USE hq_drivers_mod, ONLY: evaluate_error_on_accelerations
        INTEGER(C_INT64_T), VALUE :: n
        REAL(C_DOUBLE), INTENT(IN) :: m(n), qs(3, n), F(3, n), Fd(3, n)
        REAL(C_DOUBLE), INTENT(OUT) :: errors(7)
        CALL fd_reset()
        CALL evaluate_error_on_accelerations(m, qs, F, Fd, 1, errors)
        fd_errors = hq_status
    END FUNCTION

    ! integration.f90 drivers on (3,N) qs, ps in place; kind 0 direct, 1 BH,
    ! 2 Dehnen.  E(1:2) = total energy before and after.
    INTEGER(C_INT) FUNCTION fd_verlet(n, m, qs, ps, G, eps2, dt, tf, theta2, mult, kind, E) &
            BIND(C, name="fd_verlet")
        USE hq_drivers_mod, ONLY: test_integration_verlet_direct, test_integration_verlet_bh, &
                                  test_integration_verlet_dehnen, total_energy
        INTEGER(C_INT64_T), VALUE :: n
        REAL(C_DOUBLE), INTENT(IN) :: m(n)
        REAL(C_DOUBLE), INTENT(INOUT) :: qs(3, n), ps(3, n)
        REAL(C_DOUBLE), VALUE :: G, eps2, dt, tf, theta2
        INTEGER(C_INT), VALUE :: mult, kind
        REAL(C_DOUBLE), INTENT(OUT) :: E(2)
        CALL fd_reset()
        CALL total_energy(m, qs, ps, G, eps2, E(1))
        SELECT CASE (kind)
            CASE (0)
                CALL test_integration_verlet_direct(m, qs, ps, G, eps2, dt, tf, 1)
            CASE (1)
                CALL test_integration_verlet_bh(m, qs, ps, G, eps2, dt, tf, theta2, INT(mult), 1)
            CASE (2)
                CALL test_integration_verlet_dehnen(m, qs, ps, G, eps2, dt, tf, theta2, INT(mult))
        END SELECT
        CALL total_energy(m, qs, ps, G, eps2, E(2))
        fd_verlet = hq_status
    END FUNCTION

    ! parareal.f90:8-69: ug = G_til(m, u0, ...), uf = F_til(m, u0, ...), u0 (2N,3).
    INTEGER(C_INT) FUNCTION fd_propagators(n, m, u0, dt, qntt, theta, G, eps, ug, uf) &
            BIND(C, name="fd_propagators")
        USE hq_parareal_mod, ONLY: G_til, F_til
        INTEGER(C_INT64_T), VALUE :: n
        REAL(C_DOUBLE), INTENT(IN) :: m(n), u0(2 * n, 3)
        REAL(C_DOUBLE), VALUE :: dt, theta, G, eps
        INTEGER(C_INT), VALUE :: qntt
        REAL(C_DOUBLE), INTENT(OUT) :: ug(2 * n, 3), uf(2 * n, 3)
        CALL fd_reset()
        ug = G_til(m, u0, dt, INT(qntt), theta, G, eps)
        uf = F_til(m, u0, dt, INT(qntt), theta, G, eps)
        fd_propagators = hq_status
    END FUNCTION

    ! parareal.f90:97-186; q0, p0 (N,3); q1, p1 (nw,N,3); diag = E, P(3), J(3)
    ! of the last window's final state (parareal/api3d.f90:108-155).
    INTEGER(C_INT) FUNCTION fd_parareal(n, m, q0, p0, tf, nw, itmax, theta, G, eps, maxerr, qntt, &
                                        correction, q1, p1, iterations, errors, diag) &
            BIND(C, name="fd_parareal")
        USE hq_parareal_mod, ONLY: integrate_parareal, total_energy, total_linear_momentum, &
                                   total_angular_momentum
        INTEGER(C_INT64_T), VALUE :: n
        REAL(C_DOUBLE), INTENT(IN) :: m(n), q0(n, 3), p0(n, 3)
        REAL(C_DOUBLE), VALUE :: tf, theta, G, eps, maxerr
        INTEGER(C_INT), VALUE :: nw, itmax, qntt, correction
        REAL(C_DOUBLE), INTENT(OUT) :: q1(nw, n, 3), p1(nw, n, 3)
        INTEGER(C_INT), INTENT(OUT) :: iterations
        REAL(C_DOUBLE), INTENT(OUT) :: errors(itmax), diag(7)
        INTEGER :: it
        CALL fd_reset()
        errors = 0.0_C_DOUBLE
        CALL integrate_parareal(m, q0, p0, tf, INT(nw), INT(itmax), theta, G, eps, maxerr, INT(qntt), &
                                q1, p1, correction=INT(correction), iterations=it, errors=errors, &
                                verbose=.FALSE.)
        iterations = it
        diag(1) = total_energy(m, q1(nw, :, :), p1(nw, :, :), G, eps)
        diag(2:4) = total_linear_momentum(p1(nw, :, :))
        diag(5:7) = total_angular_momentum(q1(nw, :, :), p1(nw, :, :))
        fd_parareal = hq_status
    END FUNCTION
END MODULE
