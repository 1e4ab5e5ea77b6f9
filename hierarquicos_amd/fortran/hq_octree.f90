! hq_octree.f90 -- the reference's Fortran API, implemented on libhq.so, so that
! its Fortran drivers switch to the GPU by changing their USE lines:
!
!   reference module (file)                          this file
!   octree_mod     (fortran/octree_default.f90)       hq_octree_mod       OctreeType
!   api_mod        (fortran/api3d.f90)                hq_api_mod          generate_tree,
!                                                                         compute_forces(_direct) -> (N,3)
!   direct_forces  (fortran/direct_forces.f90)        hq_direct_mod       compute_forces_direct -> (3,N)
!   benchmarks/integration_verlet/integration.f90,
!   benchmarks/test_forces.f90                        hq_drivers_mod      test_integration_verlet_*,
!                                                                         total_energy, evaluate_error_on_accelerations
!   parareal/parareal.f90, parareal/api3d.f90         hq_parareal_mod     integrate_parareal, G_til, F_til,
!                                                                         total_energy, total_*_momentum
!
! Reals are REAL(8) throughout (the Parareal reference computes in REAL(16)).
! Errors stop with the library's message unless hq_stop_on_error is cleared
! (module hq_iface), in which case hq_status records the first failure.
MODULE hq_octree_mod
    USE ISO_C_BINDING
    USE hq_iface
    IMPLICIT NONE
    PRIVATE
    PUBLIC :: OctreeType

    INTEGER, PARAMETER :: pf = C_DOUBLE

    ! octree_default.f90:90-170.  The node store lives in HBM behind the
    ! handle; `cache` holds the forces of the last all-target evaluation that
    ! forces(p, ...) serves.  An OctreeType owns its handle: do not copy one
    ! by assignment (both copies would free it).
    TYPE :: OctreeType
        TYPE(C_PTR) :: h = C_NULL_PTR
        INTEGER :: N = 0
        INTEGER :: method = 0
        INTEGER :: group_size = 64
        INTEGER :: precision = 0
        LOGICAL :: is_allocated = .FALSE.
        INTEGER :: cached = 0
        REAL(pf) :: cache_par(3) = 0.0_pf
        REAL(pf), ALLOCATABLE :: cache(:,:)
    CONTAINS
        PROCEDURE :: pre_init, init, clear, destroy
        PROCEDURE :: forces => bh_forces_over_p
        PROCEDURE :: forces_all, dehnen_eval, forces_groups, forces_bunched, fmm
        PROCEDURE :: tune, autotune, number_of_nodes
        FINAL :: octree_final
    END TYPE

CONTAINS

    ! pre_init (octree_default.f90:178-282): method 10/11/12 (BH mono, quad,
    ! oct), 20/21/22 (Dehnen), 30/31/32 (grouped BH).  precision: HQ_FP64
    ! (default) or HQ_MIXED.  Collision detection (collide_par, radii) is not
    ! part of this library.
    SUBROUTINE pre_init(self, m, method_par, collide_par, radii, group_size, precision)
        CLASS(OctreeType), INTENT(INOUT) :: self
        REAL(pf), INTENT(IN) :: m(:)
        INTEGER, INTENT(IN), OPTIONAL :: method_par
        LOGICAL, INTENT(IN), OPTIONAL :: collide_par
        REAL(pf), INTENT(IN), OPTIONAL :: radii(:)
        INTEGER, INTENT(IN), OPTIONAL :: group_size
        INTEGER, INTENT(IN), OPTIONAL :: precision

        IF (PRESENT(collide_par)) THEN
            IF (collide_par) THEN
                WRITE(*, '(A)') 'pre_init: collision detection is not supported'
                ERROR STOP 1
            END IF
        END IF
        IF (PRESENT(radii)) CONTINUE
        CALL self % destroy()
        self % method = 0
        IF (PRESENT(method_par)) self % method = method_par
        self % group_size = 64
        IF (PRESENT(group_size)) self % group_size = group_size
        self % precision = HQ_FP64
        IF (PRESENT(precision)) self % precision = precision
        self % N = SIZE(m)
        CALL hq_check(hq_tree_create(self % h, INT(self % N, C_INT64_T), &
                                     INT(self % method, C_INT), INT(self % group_size, C_INT), &
                                     INT(self % precision, C_INT)), 'pre_init')
        IF (.NOT. C_ASSOCIATED(self % h)) RETURN
        CALL hq_check(hq_tree_set_masses(self % h, m, HQ_HOST), 'pre_init')
        self % is_allocated = .TRUE.
    END SUBROUTINE

    ! init (:320-387): keys, sort, build, upward pass, groups.  save_txt (a
    ! text dump of the root) is accepted and ignored.
    SUBROUTINE init(self, x, y, z, save_txt)
        CLASS(OctreeType), INTENT(INOUT) :: self
        REAL(pf), INTENT(IN) :: x(:), y(:), z(:)
        INTEGER, INTENT(IN), OPTIONAL :: save_txt
        IF (PRESENT(save_txt)) CONTINUE
        self % cached = 0
        IF (SIZE(x) /= self % N .OR. SIZE(y) /= self % N .OR. SIZE(z) /= self % N) THEN
            CALL hq_check(HQ_ERR_ARGS, 'init: x, y, z must have the length of m')
            RETURN
        END IF
        CALL hq_check(hq_tree_build(self % h, x, y, z, HQ_HOST), 'init')
    END SUBROUTINE

    ! clear (:284-318) zeroes the node store before each init; the GPU build
    ! overwrites it, so there is nothing to do.
    SUBROUTINE clear(self)
        CLASS(OctreeType), INTENT(INOUT) :: self
        self % cached = 0
    END SUBROUTINE

    ! Frees the device store (also done by finalisation).
    SUBROUTINE destroy(self)
        CLASS(OctreeType), INTENT(INOUT) :: self
        INTEGER(C_INT) :: st
        IF (C_ASSOCIATED(self % h)) st = hq_tree_destroy(self % h)
        self % h = C_NULL_PTR
        self % is_allocated = .FALSE.
        self % cached = 0
        IF (ALLOCATED(self % cache)) DEALLOCATE(self % cache)
    END SUBROUTINE

    SUBROUTINE octree_final(self)
        TYPE(OctreeType), INTENT(INOUT) :: self
        CALL self % destroy()
    END SUBROUTINE

    ! bh_forces_over_p (:895-1078): the force on particle p (1-based) with
    ! theta^2, G and eps^2 (defaults 0, 1, 0).  The GPU evaluates every target
    ! on the first call after init and later calls with the same parameters
    ! read that result, so the reference's loops over p (also under OpenMP,
    ! test_forces.f90:199-213) cost one walk.  Methods 10-12.
    FUNCTION bh_forces_over_p(self, p, par_theta2, par_G, par_eps2) RESULT(forces)
        CLASS(OctreeType), INTENT(INOUT) :: self
        INTEGER, INTENT(IN) :: p
        REAL(pf), INTENT(IN), OPTIONAL :: par_theta2, par_G, par_eps2
        REAL(pf) :: forces(3)
        REAL(pf) :: par(3)

        par = [0.0_pf, 1.0_pf, 0.0_pf]
        IF (PRESENT(par_theta2)) par(1) = par_theta2
        IF (PRESENT(par_G)) par(2) = par_G
        IF (PRESENT(par_eps2)) par(3) = par_eps2
        forces = 0.0_pf
        IF (self % method < 10 .OR. self % method > 12) THEN
            CALL hq_check(HQ_ERR_ARGS, 'forces: needs method 10, 11 or 12')
            RETURN
        END IF
        ! Double-checked fill: the flag is published with release order after
        ! the cache is written and read with acquire order before it is used.
        IF (.NOT. cache_ok(self, par)) THEN
            !$OMP CRITICAL (hq_octree_cache)
            IF (.NOT. cache_ok(self, par)) THEN
                IF (.NOT. ALLOCATED(self % cache)) ALLOCATE(self % cache(3, self % N))
                CALL hq_check(hq_tree_forces(self % h, par(1), par(2), par(3), self % cache, &
                                             HQ_LAYOUT_3N, HQ_HOST), 'forces')
                self % cache_par = par
                !$OMP ATOMIC WRITE RELEASE
                self % cached = 1
            END IF
            !$OMP END CRITICAL (hq_octree_cache)
        END IF
        IF (p >= 1 .AND. p <= self % N) forces = self % cache(:, p)
    END FUNCTION

    LOGICAL FUNCTION cache_ok(self, par)
        CLASS(OctreeType), INTENT(IN) :: self
        REAL(pf), INTENT(IN) :: par(3)
        INTEGER :: c
        !$OMP ATOMIC READ ACQUIRE
        c = self % cached
        cache_ok = c == 1
        IF (cache_ok) cache_ok = ALL(self % cache_par == par)
    END FUNCTION

    ! Every target at once with the handle's own method, forces(3,N).
    SUBROUTINE forces_all(self, theta2, G, eps2, forces)
        CLASS(OctreeType), INTENT(INOUT) :: self
        REAL(pf), INTENT(IN) :: theta2, G, eps2
        REAL(pf), INTENT(OUT) :: forces(:,:)
        IF (SIZE(forces) < 3 * self % N) THEN
            CALL hq_check(HQ_ERR_ARGS, 'forces_all: forces must be (3,N)')
            RETURN
        END IF
        CALL hq_check(hq_tree_forces(self % h, theta2, G, eps2, forces, HQ_LAYOUT_3N, HQ_HOST), &
                      'forces_all')
    END SUBROUTINE

    ! dehnen_eval (:1680-1712), methods 20-22; nt_par has no GPU meaning.
    SUBROUTINE dehnen_eval(self, theta2, potsoft2, G, forces, nt_par)
        CLASS(OctreeType), INTENT(INOUT) :: self
        REAL(pf), INTENT(IN) :: theta2, potsoft2, G
        REAL(pf), INTENT(OUT) :: forces(:,:)
        INTEGER, INTENT(IN), OPTIONAL :: nt_par
        IF (PRESENT(nt_par)) CONTINUE
        IF (self % method < 20 .OR. self % method > 22) THEN
            CALL hq_check(HQ_ERR_ARGS, 'dehnen_eval: needs method 20, 21 or 22')
            RETURN
        END IF
        CALL self % forces_all(theta2, G, potsoft2, forces)
    END SUBROUTINE

    ! forces_groups (:1811-2095), methods 30-32; nt_par has no GPU meaning.
    SUBROUTINE forces_groups(self, theta2, G, eps2, forces, nt_par)
        CLASS(OctreeType), INTENT(INOUT) :: self
        REAL(pf), INTENT(IN) :: theta2, G, eps2
        REAL(pf), INTENT(OUT) :: forces(:,:)
        INTEGER, INTENT(IN), OPTIONAL :: nt_par
        IF (PRESENT(nt_par)) CONTINUE
        IF (self % method < 30 .OR. self % method > 32) THEN
            CALL hq_check(HQ_ERR_ARGS, 'forces_groups: needs method 30, 31 or 32')
            RETURN
        END IF
        CALL self % forces_all(theta2, G, eps2, forces)
    END SUBROUTINE

    ! The bunched walk (no reference counterpart): 64-body bunches under the
    ! grouped acceptance test, FP32 interactions when precision = HQ_MIXED.
    SUBROUTINE forces_bunched(self, theta2, G, eps2, forces, precision)
        CLASS(OctreeType), INTENT(INOUT) :: self
        REAL(pf), INTENT(IN) :: theta2, G, eps2
        REAL(pf), INTENT(OUT) :: forces(:,:)
        INTEGER, INTENT(IN), OPTIONAL :: precision
        INTEGER(C_INT) :: prec
        prec = HQ_MIXED
        IF (PRESENT(precision)) prec = INT(precision, C_INT)
        IF (SIZE(forces) < 3 * self % N) THEN
            CALL hq_check(HQ_ERR_ARGS, 'forces_bunched: forces must be (3,N)')
            RETURN
        END IF
        CALL hq_check(hq_tree_forces_bunched(self % h, theta2, G, eps2, forces, HQ_LAYOUT_3N, &
                                             HQ_HOST, prec), 'forces_bunched')
    END SUBROUTINE

    ! Cartesian FMM of order 4 or 6 (mode HQ_FMM_MFMA, the default, or HQ_FMM_FP64).
    SUBROUTINE fmm(self, order, theta2, G, eps2, forces, mode)
        CLASS(OctreeType), INTENT(INOUT) :: self
        INTEGER, INTENT(IN) :: order
        REAL(pf), INTENT(IN) :: theta2, G, eps2
        REAL(pf), INTENT(OUT) :: forces(:,:)
        INTEGER, INTENT(IN), OPTIONAL :: mode
        INTEGER(C_INT) :: md
        md = HQ_FMM_MFMA
        IF (PRESENT(mode)) md = INT(mode, C_INT)
        IF (SIZE(forces) < 3 * self % N) THEN
            CALL hq_check(HQ_ERR_ARGS, 'fmm: forces must be (3,N)')
            RETURN
        END IF
        CALL hq_check(hq_tree_fmm(self % h, INT(order, C_INT), theta2, G, eps2, md, forces, &
                                  HQ_LAYOUT_3N, HQ_HOST), 'fmm')
    END SUBROUTINE

    ! One knob of the bunched walk or the FMM (HQ_TUNE_*).
    SUBROUTINE tune(self, key, value)
        CLASS(OctreeType), INTENT(INOUT) :: self
        INTEGER, INTENT(IN) :: key, value
        self % cached = 0
        CALL hq_check(hq_tree_tune(self % h, INT(key, C_INT), INT(value, C_INT64_T)), 'tune')
    END SUBROUTINE

    ! Times the bunched walk over its knobs on the built tree and keeps the fastest.
    SUBROUTINE autotune(self, theta2, G, eps2, precision, bunch, super_size, leaf, ms)
        CLASS(OctreeType), INTENT(INOUT) :: self
        REAL(pf), INTENT(IN) :: theta2, G, eps2
        INTEGER, INTENT(IN) :: precision
        INTEGER, INTENT(OUT) :: bunch, super_size, leaf
        REAL(pf), INTENT(OUT) :: ms
        INTEGER(C_INT) :: b, s, l
        CALL hq_check(hq_tree_autotune(self % h, theta2, G, eps2, INT(precision, C_INT), 2_C_INT, &
                                       b, s, l, ms), 'autotune')
        bunch = b
        super_size = s
        leaf = l
    END SUBROUTINE

    INTEGER FUNCTION number_of_nodes(self)
        CLASS(OctreeType), INTENT(IN) :: self
        TYPE(hq_tree_info_t) :: info
        number_of_nodes = 0
        CALL hq_check(hq_tree_info(self % h, info), 'number_of_nodes')
        number_of_nodes = INT(info % nnodes)
    END FUNCTION
END MODULE


! api3d.f90:9-72, the f2py module's three entry points.  theta is the opening
! angle and eps the softening length; quad selects method 11 over 10, method
! any of 10-12, 20-22, 30-32, and legacy_theta2 the shipped call's reading of
! theta as theta^2 (api3d.f90:37).
MODULE hq_api_mod
    USE ISO_C_BINDING
    USE hq_iface
    IMPLICIT NONE
    PRIVATE
    PUBLIC :: generate_tree, compute_forces, compute_forces_direct
    INTEGER, PARAMETER :: pf = C_DOUBLE
CONTAINS

    INTEGER(C_INT) FUNCTION api_method(quad, method)
        LOGICAL, INTENT(IN), OPTIONAL :: quad
        INTEGER, INTENT(IN), OPTIONAL :: method
        api_method = 10
        IF (PRESENT(quad)) THEN
            IF (quad) api_method = 11
        END IF
        IF (PRESENT(method)) api_method = INT(method, C_INT)
    END FUNCTION

    SUBROUTINE generate_tree(m, x, y, z, quad, method)
        REAL(pf), INTENT(IN) :: m(:), x(:), y(:), z(:)
        LOGICAL, INTENT(IN), OPTIONAL :: quad
        INTEGER, INTENT(IN), OPTIONAL :: method
        CALL hq_check(hq_generate_tree(INT(SIZE(m), C_INT64_T), m, x, y, z, &
                                       api_method(quad, method)), 'generate_tree')
    END SUBROUTINE

    FUNCTION compute_forces(m, x, y, z, theta, G, eps, quad, method, legacy_theta2) RESULT(forces)
        REAL(pf), INTENT(IN) :: m(:), x(:), y(:), z(:), theta, G, eps
        LOGICAL, INTENT(IN), OPTIONAL :: quad, legacy_theta2
        INTEGER, INTENT(IN), OPTIONAL :: method
        REAL(pf) :: forces(SIZE(m), 3)
        INTEGER(C_INT) :: legacy
        legacy = 0
        IF (PRESENT(legacy_theta2)) THEN
            IF (legacy_theta2) legacy = 1
        END IF
        forces = 0.0_pf
        CALL hq_check(hq_compute_forces(INT(SIZE(m), C_INT64_T), m, x, y, z, theta, G, eps, &
                                        api_method(quad, method), legacy, forces), 'compute_forces')
    END FUNCTION

    FUNCTION compute_forces_direct(m, x, y, z, G, eps) RESULT(forces)
        REAL(pf), INTENT(IN) :: m(:), x(:), y(:), z(:), G, eps
        REAL(pf) :: forces(SIZE(m), 3)
        forces = 0.0_pf
        CALL hq_check(hq_compute_forces_direct(INT(SIZE(m), C_INT64_T), m, x, y, z, G, eps, &
                                               forces), 'compute_forces_direct')
    END FUNCTION
END MODULE


! fortran/direct_forces.f90:8-64: compute_forces_direct(m, x, y, z, G, eps2, nt)
! -> forces(3,N), all pairs in FP64 (eps2 is the squared softening; nt has no
! GPU meaning).
MODULE hq_direct_mod
    USE ISO_C_BINDING
    USE hq_iface
    IMPLICIT NONE
    PRIVATE
    PUBLIC :: compute_forces_direct
    INTEGER, PARAMETER :: pf = C_DOUBLE
CONTAINS

    FUNCTION compute_forces_direct(m, x, y, z, G, eps2, nt) RESULT(forces)
        REAL(pf), INTENT(IN), TARGET :: m(:), x(:), y(:), z(:)
        REAL(pf), INTENT(IN) :: G, eps2
        INTEGER, INTENT(IN), OPTIONAL :: nt
        REAL(pf), TARGET :: forces(3, SIZE(m))
        TYPE(C_PTR) :: d(5)
        INTEGER(C_INT64_T) :: n, bytes
        INTEGER(C_INT) :: st
        INTEGER :: k
        IF (PRESENT(nt)) CONTINUE
        forces = 0.0_pf
        n = SIZE(m)
        IF (n == 0) RETURN
        IF (SIZE(x) /= n .OR. SIZE(y) /= n .OR. SIZE(z) /= n) THEN
            CALL hq_check(HQ_ERR_ARGS, 'compute_forces_direct: x, y, z must have the length of m')
            RETURN
        END IF
        ! m, x, y, z and F in HBM; the gather-form direct sum over all targets.
        bytes = 8_C_INT64_T * n
        d = C_NULL_PTR
        st = HQ_OK
        DO k = 1, 5
            IF (st == HQ_OK) st = hq_dev_alloc(d(k), MERGE(3 * bytes, bytes, k == 5))
        END DO
        IF (st == HQ_OK) st = hq_dev_copy(d(1), C_LOC(m), bytes, 0_C_INT)
        IF (st == HQ_OK) st = hq_dev_copy(d(2), C_LOC(x), bytes, 0_C_INT)
        IF (st == HQ_OK) st = hq_dev_copy(d(3), C_LOC(y), bytes, 0_C_INT)
        IF (st == HQ_OK) st = hq_dev_copy(d(4), C_LOC(z), bytes, 0_C_INT)
        IF (st == HQ_OK) st = hq_direct_forces_device(n, d(1), d(2), d(3), d(4), n, C_NULL_PTR, &
                                                      G, eps2, HQ_FP64, d(5))
        IF (st == HQ_OK) st = hq_dev_copy(C_LOC(forces), d(5), 3 * bytes, 1_C_INT)
        CALL hq_check(st, 'compute_forces_direct')
        DO k = 1, 5
            IF (C_ASSOCIATED(d(k))) st = hq_dev_free(d(k))
        END DO
    END FUNCTION
END MODULE


! The drivers of benchmarks/integration_verlet/integration.f90:85-230 and the
! harness metric of benchmarks/test_forces.f90:66-127.  qs, ps and forces are
! (3,N); the state stays in HBM for the whole integration and the tree is
! rebuilt before every force evaluation, as the reference's loops call init.
MODULE hq_drivers_mod
    USE ISO_C_BINDING
    USE hq_iface
    IMPLICIT NONE
    PRIVATE
    PUBLIC :: test_integration_verlet_direct, test_integration_verlet_bh
    PUBLIC :: test_integration_verlet_dehnen, total_energy, evaluate_error_on_accelerations
    PUBLIC :: steps_until
    INTEGER, PARAMETER :: pf = C_DOUBLE
CONTAINS

    ! Step count of DO WHILE (t < tf); ...; t = t + ABS(dt) (integration.f90:150-171).
    INTEGER(C_INT64_T) FUNCTION steps_until(tf, dt) RESULT(n)
        REAL(pf), INTENT(IN) :: tf, dt
        REAL(pf) :: t
        t = 0.0_pf
        n = 0
        DO WHILE (t < tf)
            t = t + ABS(dt)
            n = n + 1
        END DO
    END FUNCTION

    ! A handle of `method` holding (ms, qs, ps); nsteps of `scheme` with `force`;
    ! the final state back into qs, ps.
    SUBROUTINE run_state(ms, qs, ps, method, scheme, force, dt, nsteps, theta2, G, eps2, what)
        REAL(pf), INTENT(IN) :: ms(:)
        REAL(pf), INTENT(INOUT) :: qs(:,:), ps(:,:)
        INTEGER(C_INT), INTENT(IN) :: method, scheme, force
        REAL(pf), INTENT(IN) :: dt, theta2, G, eps2
        INTEGER(C_INT64_T), INTENT(IN) :: nsteps
        CHARACTER(LEN=*), INTENT(IN) :: what
        TYPE(C_PTR) :: h
        INTEGER(C_INT) :: st, st2
        REAL(pf), ALLOCATABLE :: c(:,:)
        INTEGER :: n
        n = SIZE(ms)
        IF (SIZE(qs, 2) /= n .OR. SIZE(ps, 2) /= n .OR. SIZE(qs, 1) /= 3 .OR. SIZE(ps, 1) /= 3) THEN
            CALL hq_check(HQ_ERR_ARGS, what // ': qs and ps must be (3,N)')
            RETURN
        END IF
        ALLOCATE(c(n, 6))
        c(:, 1:3) = TRANSPOSE(qs)
        c(:, 4:6) = TRANSPOSE(ps)
        h = C_NULL_PTR
        st = hq_tree_create(h, INT(n, C_INT64_T), method, 64_C_INT, HQ_FP64)
        IF (st == HQ_OK) st = hq_tree_set_masses(h, ms, HQ_HOST)
        IF (st == HQ_OK) st = hq_tree_set_state(h, c(:,1), c(:,2), c(:,3), c(:,4), c(:,5), &
                                                c(:,6), HQ_HOST)
        IF (st == HQ_OK) st = hq_tree_integrate(h, scheme, dt, nsteps, theta2, G, eps2, force)
        IF (st == HQ_OK) st = hq_tree_get_state(h, c(:,1), c(:,2), c(:,3), c(:,4), c(:,5), &
                                                c(:,6), HQ_HOST)
        IF (C_ASSOCIATED(h)) st2 = hq_tree_destroy(h)
        CALL hq_check(st, what)
        IF (st /= HQ_OK) RETURN
        qs = TRANSPOSE(c(:, 1:3))
        ps = TRANSPOSE(c(:, 4:6))
    END SUBROUTINE

    ! integration.f90:85-108: Verlet with the FP64 direct sum.
    SUBROUTINE test_integration_verlet_direct(ms, qs, ps, G, eps2, dt, tf, nt)
        REAL(pf), INTENT(INOUT) :: qs(:,:), ps(:,:)
        REAL(pf), INTENT(IN) :: ms(:), G, dt, tf, eps2
        INTEGER, INTENT(IN) :: nt
        IF (nt < 0) CONTINUE
        CALL run_state(ms, qs, ps, 10_C_INT, HQ_SCHEME_VERLET, HQ_FORCE_DIRECT, dt, &
                       steps_until(tf, dt), 0.0_pf, G, eps2, 'test_integration_verlet_direct')
    END SUBROUTINE

    ! integration.f90:110-172: Verlet with BH forces of order mult (1, 4, 8).
    SUBROUTINE test_integration_verlet_bh(ms, qs, ps, G, eps2, dt, tf, theta2, mult, nt)
        REAL(pf), INTENT(INOUT) :: qs(:,:), ps(:,:)
        REAL(pf), INTENT(IN) :: ms(:), G, dt, tf, eps2, theta2
        INTEGER, INTENT(IN) :: nt, mult
        INTEGER(C_INT) :: method
        IF (nt < 0) CONTINUE
        method = 10
        IF (mult == 4) method = 11
        IF (mult == 8) method = 12
        CALL run_state(ms, qs, ps, method, HQ_SCHEME_VERLET, HQ_FORCE_TREE, dt, &
                       steps_until(tf, dt), theta2, G, eps2, 'test_integration_verlet_bh')
    END SUBROUTINE

    ! integration.f90:174-209: Verlet with Dehnen forces (mult 1 or 4).
    SUBROUTINE test_integration_verlet_dehnen(ms, qs, ps, G, eps2, dt, tf, theta2, mult)
        REAL(pf), INTENT(INOUT) :: qs(:,:), ps(:,:)
        REAL(pf), INTENT(IN) :: ms(:), G, dt, tf, eps2, theta2
        INTEGER, INTENT(IN) :: mult
        INTEGER(C_INT) :: method
        method = 20
        IF (mult == 4) method = 21
        CALL run_state(ms, qs, ps, method, HQ_SCHEME_VERLET, HQ_FORCE_TREE, dt, &
                       steps_until(tf, dt), theta2, G, eps2, 'test_integration_verlet_dehnen')
    END SUBROUTINE

    ! integration.f90:211-230: kinetic plus softened pair potential energy.
    SUBROUTINE total_energy(ms, qs, ps, G, eps2, E)
        REAL(pf), INTENT(IN) :: ms(:), qs(:,:), ps(:,:), G, eps2
        REAL(pf), INTENT(OUT) :: E
        TYPE(C_PTR) :: h
        INTEGER(C_INT) :: st, st2
        REAL(pf), ALLOCATABLE :: c(:,:)
        REAL(pf) :: out(7)
        INTEGER :: n
        n = SIZE(ms)
        E = 0.0_pf
        ALLOCATE(c(n, 6))
        c(:, 1:3) = TRANSPOSE(qs)
        c(:, 4:6) = TRANSPOSE(ps)
        h = C_NULL_PTR
        st = hq_tree_create(h, INT(n, C_INT64_T), 10_C_INT, 64_C_INT, HQ_FP64)
        IF (st == HQ_OK) st = hq_tree_set_masses(h, ms, HQ_HOST)
        IF (st == HQ_OK) st = hq_tree_set_state(h, c(:,1), c(:,2), c(:,3), c(:,4), c(:,5), &
                                                c(:,6), HQ_HOST)
        IF (st == HQ_OK) st = hq_tree_energy(h, G, eps2, out)
        IF (C_ASSOCIATED(h)) st2 = hq_tree_destroy(h)
        CALL hq_check(st, 'total_energy')
        IF (st == HQ_OK) E = out(1)
    END SUBROUTINE

    ! test_forces.f90:66-127: errors(1) McMillan-Aarseth RMS acceleration error,
    ! errors(2:4) the sum of the forces, errors(5:7) the total torque.
    SUBROUTINE evaluate_error_on_accelerations(ms, qs, forces, forces_dir, nt, errors)
        REAL(pf), INTENT(IN) :: ms(:), qs(:,:)
        REAL(pf), INTENT(OUT) :: errors(:)
        REAL(pf), INTENT(IN) :: forces(:,:), forces_dir(:,:)
        INTEGER, INTENT(IN) :: nt
        REAL(pf) :: e(7)
        IF (nt < 0) CONTINUE
        errors = 0.0_pf
        CALL hq_check(hq_errors_ma(INT(SIZE(ms), C_INT64_T), ms, qs(1,:), qs(2,:), qs(3,:), forces, &
                                   forces_dir, HQ_LAYOUT_3N, HQ_HOST, e), &
                      'evaluate_error_on_accelerations')
        errors(1:MIN(7, SIZE(errors))) = e(1:MIN(7, SIZE(errors)))
    END SUBROUTINE
END MODULE


! parareal/parareal.f90:8-186 and the diagnostics of parareal/api3d.f90:108-155,
! in REAL(8).  States are (N,3) and u0 is (2N,3) (positions over momenta).
MODULE hq_parareal_mod
    USE ISO_C_BINDING
    USE hq_iface
    IMPLICIT NONE
    PRIVATE
    PUBLIC :: integrate_parareal, G_til, F_til
    PUBLIC :: total_energy, total_linear_momentum, total_angular_momentum
    INTEGER, PARAMETER :: pf = C_DOUBLE
CONTAINS

    ! qntt steps of `scheme` from u0 with `force` on a handle of `method`.
    FUNCTION propagate(m, u0, dt, qntt, method, scheme, force, theta2, G, eps2, what) RESULT(u1)
        REAL(pf), INTENT(IN) :: m(:), u0(:,:), dt, theta2, G, eps2
        INTEGER, INTENT(IN) :: qntt
        INTEGER(C_INT), INTENT(IN) :: method, scheme, force
        CHARACTER(LEN=*), INTENT(IN) :: what
        REAL(pf) :: u1(SIZE(u0, 1), 3)
        TYPE(C_PTR) :: h
        INTEGER(C_INT) :: st, st2
        INTEGER :: n
        n = SIZE(m)
        u1 = u0
        IF (SIZE(u0, 1) /= 2 * n .OR. SIZE(u0, 2) /= 3) THEN
            CALL hq_check(HQ_ERR_ARGS, what // ': u0 must be (2N,3)')
            RETURN
        END IF
        h = C_NULL_PTR
        st = hq_tree_create(h, INT(n, C_INT64_T), method, 64_C_INT, HQ_FP64)
        IF (st == HQ_OK) st = hq_tree_set_masses(h, m, HQ_HOST)
        IF (st == HQ_OK) st = hq_tree_set_state(h, u0(1:n,1), u0(1:n,2), u0(1:n,3), &
                                                u0(n+1:,1), u0(n+1:,2), u0(n+1:,3), HQ_HOST)
        IF (st == HQ_OK) st = hq_tree_integrate(h, scheme, dt, INT(qntt, C_INT64_T), theta2, G, &
                                                eps2, force)
        IF (st == HQ_OK) st = hq_tree_get_state(h, u1(1:n,1), u1(1:n,2), u1(1:n,3), &
                                                u1(n+1:,1), u1(n+1:,2), u1(n+1:,3), HQ_HOST)
        IF (C_ASSOCIATED(h)) st2 = hq_tree_destroy(h)
        CALL hq_check(st, what)
    END FUNCTION

    ! parareal.f90:8-44: qntt symplectic Euler steps with quadrupole BH forces
    ! (theta squared, parareal/api3d.f90:30) or, for theta == -1, direct forces.
    FUNCTION G_til(m, u0, dt, qntt, theta, G, eps) RESULT(u1)
        REAL(pf), INTENT(IN) :: m(:), u0(:,:), dt, theta, G, eps
        INTEGER, INTENT(IN) :: qntt
        REAL(pf) :: u1(SIZE(u0, 1), 3)
        IF (theta == -1.0_pf) THEN
            u1 = propagate(m, u0, dt, qntt, 11_C_INT, HQ_SCHEME_EULER, HQ_FORCE_DIRECT, 0.0_pf, &
                           G, eps * eps, 'G_til')
        ELSE
            u1 = propagate(m, u0, dt, qntt, 11_C_INT, HQ_SCHEME_EULER, HQ_FORCE_TREE, theta * theta, &
                           G, eps * eps, 'G_til')
        END IF
    END FUNCTION

    ! parareal.f90:46-69: qntt Verlet steps with direct forces.
    FUNCTION F_til(m, u0, dt, qntt, theta, G, eps) RESULT(u1)
        REAL(pf), INTENT(IN) :: m(:), u0(:,:), dt, theta, G, eps
        INTEGER, INTENT(IN) :: qntt
        REAL(pf) :: u1(SIZE(u0, 1), 3)
        IF (theta < 0.0_pf) CONTINUE
        u1 = propagate(m, u0, dt, qntt, 10_C_INT, HQ_SCHEME_VERLET, HQ_FORCE_DIRECT, 0.0_pf, G, &
                       eps * eps, 'F_til')
    END FUNCTION

    ! parareal.f90:97-186 with the reference's propagators: coarse = one Euler
    ! step per window with quadrupole BH (direct for theta == -1), fine = qntt
    ! Verlet steps with direct forces; all fine windows of an iteration run as
    ! one batch.  q0, p0 are (N,3); q1, p1 (num_windows,N,3) hold the last
    ! iterate.  correction: HQ_CORR_EXACT (default: a converged window equals
    ! the sequential fine solution bit for bit) or HQ_CORR_REFERENCE (the
    ! reference's (G_new + F) - G_old).  iterations and errors (the error of
    ! iterates 2, 3, ...) are optional outputs; verbose (default .TRUE.) prints
    ! the reference's per-iteration and summary lines.
    SUBROUTINE integrate_parareal(m, q0, p0, tf, num_windows, N_itermax, theta, G, eps, maxerror, &
                                  qntt, q1, p1, correction, iterations, errors, verbose)
        REAL(pf), INTENT(IN) :: m(:), q0(:,:), p0(:,:), theta, G, eps
        REAL(pf), INTENT(IN) :: tf, maxerror
        INTEGER, INTENT(IN) :: num_windows, N_itermax, qntt
        REAL(pf), INTENT(OUT) :: q1(:,:,:), p1(:,:,:)
        INTEGER, INTENT(IN), OPTIONAL :: correction
        INTEGER, INTENT(OUT), OPTIONAL :: iterations
        REAL(pf), INTENT(OUT), OPTIONAL :: errors(:)
        LOGICAL, INTENT(IN), OPTIONAL :: verbose
        TYPE(hq_parareal_params) :: prm
        REAL(pf), ALLOCATABLE :: U(:,:,:), err(:)
        INTEGER(C_INT) :: it, st
        INTEGER :: n, k
        LOGICAL :: talk

        n = SIZE(m)
        q1 = 0.0_pf
        p1 = 0.0_pf
        talk = .TRUE.
        IF (PRESENT(verbose)) talk = verbose
        IF (SIZE(q1, 1) /= num_windows .OR. SIZE(q1, 2) /= n .OR. SIZE(p1, 1) /= num_windows &
            .OR. SIZE(p1, 2) /= n) THEN
            CALL hq_check(HQ_ERR_ARGS, 'integrate_parareal: q1 and p1 must be (num_windows,N,3)')
            RETURN
        END IF
        prm % n = n
        prm % num_windows = num_windows
        prm % itermax = N_itermax
        prm % tf = tf
        prm % maxerror = maxerror
        prm % G = G
        prm % eps2 = eps * eps
        IF (theta == -1.0_pf) THEN
            prm % coarse_force = HQ_FORCE_DIRECT
        ELSE
            prm % coarse_force = HQ_FORCE_TREE
        END IF
        prm % coarse_method = 11
        prm % coarse_precision = HQ_FP64
        prm % coarse_steps = 1
        prm % coarse_theta2 = theta * theta
        prm % fine_force = HQ_FORCE_DIRECT
        prm % fine_method = 10
        prm % fine_precision = HQ_FP64
        prm % fine_steps = qntt
        prm % fine_theta2 = 0.0_pf
        prm % group_size = 64
        prm % correction = HQ_CORR_EXACT
        IF (PRESENT(correction)) prm % correction = INT(correction, C_INT)
        prm % coarse_scheme = HQ_SCHEME_EULER
        ! U(a, w, c), c = x, y, z, px, py, pz: the library's component-major layout
        ALLOCATE(U(n, num_windows, 6), err(MAX(N_itermax - 1, 1)))
        err = 0.0_pf
        it = 0
        st = hq_parareal(prm, m, q0(:,1), q0(:,2), q0(:,3), p0(:,1), p0(:,2), p0(:,3), HQ_HOST, U, &
                         HQ_HOST, err, it)
        CALL hq_check(st, 'integrate_parareal')
        IF (st /= HQ_OK) RETURN
        DO k = 1, 3
            q1(:, :, k) = TRANSPOSE(U(:, :, k))
            p1(:, :, k) = TRANSPOSE(U(:, :, k + 3))
        END DO
        IF (PRESENT(iterations)) iterations = it
        IF (PRESENT(errors)) THEN
            errors = 0.0_pf
            k = MIN(SIZE(errors), MAX(it - 1, 0))
            errors(1:k) = err(1:k)
        END IF
        IF (talk) THEN
            DO k = 2, it
                WRITE(*, '(A6,I4,A9,ES12.4)') "Iter.:", k, " / Error:", err(k - 1)
            END DO
            WRITE(*, *)
            WRITE(*, '(A17,ES12.4)') 'Max Error Accep.', maxerror
            IF (it > 1) WRITE(*, '(A17,ES12.4)') 'Final error', err(it - 1)
        END IF
    END SUBROUTINE

    ! parareal/api3d.f90:108-122 (eps is the softening length).
    FUNCTION total_energy(m, qs, ps, G, eps) RESULT(E)
        REAL(pf), INTENT(IN) :: m(:), qs(:,:), ps(:,:), G, eps
        REAL(pf) :: E
        TYPE(C_PTR) :: h
        INTEGER(C_INT) :: st, st2
        REAL(pf) :: out(7)
        INTEGER :: n
        n = SIZE(m)
        E = 0.0_pf
        h = C_NULL_PTR
        st = hq_tree_create(h, INT(n, C_INT64_T), 10_C_INT, 64_C_INT, HQ_FP64)
        IF (st == HQ_OK) st = hq_tree_set_masses(h, m, HQ_HOST)
        IF (st == HQ_OK) st = hq_tree_set_state(h, qs(:,1), qs(:,2), qs(:,3), ps(:,1), ps(:,2), &
                                                ps(:,3), HQ_HOST)
        IF (st == HQ_OK) st = hq_tree_energy(h, G, eps * eps, out)
        IF (C_ASSOCIATED(h)) st2 = hq_tree_destroy(h)
        CALL hq_check(st, 'total_energy')
        IF (st == HQ_OK) E = out(1)
    END FUNCTION

    ! parareal/api3d.f90:124-133
    FUNCTION total_linear_momentum(ps) RESULT(P)
        REAL(pf), INTENT(IN) :: ps(:,:)
        REAL(pf) :: P(3)
        INTEGER :: a
        P = 0.0_pf
        DO a = 1, SIZE(ps, 1)
            P = P + ps(a, :)
        END DO
    END FUNCTION

    ! parareal/api3d.f90:144-153
    FUNCTION total_angular_momentum(qs, ps) RESULT(J)
        REAL(pf), INTENT(IN) :: qs(:,:), ps(:,:)
        REAL(pf) :: J(3)
        INTEGER :: a
        J = 0.0_pf
        DO a = 1, SIZE(ps, 1)
            J(1) = J(1) + (qs(a,2) * ps(a,3) - ps(a,2) * qs(a,3))
            J(2) = J(2) + (-qs(a,1) * ps(a,3) + ps(a,1) * qs(a,3))
            J(3) = J(3) + (qs(a,1) * ps(a,2) - ps(a,1) * qs(a,2))
        END DO
    END FUNCTION
END MODULE
