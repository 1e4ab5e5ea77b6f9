! hq_iface.f90 -- ISO_C_BINDING interface to every entry point of libhq.so
! (include/hq.h), for Fortran drivers.  The modules of hq_octree.f90 build the
! reference's own Fortran API (OctreeType, api_mod, compute_forces_direct, the
! Verlet and Parareal drivers) on top of these declarations.
!
! Array arguments are host arrays (where = HQ_HOST) unless the argument is a
! TYPE(C_PTR): those take device pointers (hq_dev_alloc) or, where include/hq.h
! allows NULL, C_NULL_PTR.  Link with -L<dir of libhq.so> -lhq.
MODULE hq_iface
    USE ISO_C_BINDING
    IMPLICIT NONE

    ! status codes
    INTEGER(C_INT), PARAMETER :: HQ_OK = 0, HQ_ERR_ARGS = 1, HQ_ERR_CAPACITY = 2
    INTEGER(C_INT), PARAMETER :: HQ_ERR_LOST = 3, HQ_ERR_HIP = 4, HQ_ERR_NONFINITE = 5
    INTEGER(C_INT), PARAMETER :: HQ_ERR_DEPTH = 6, HQ_ERR_NOMEM = 7, HQ_ERR_OVERFLOW = 8
    ! memory, precision, layouts
    INTEGER(C_INT), PARAMETER :: HQ_HOST = 0, HQ_DEVICE = 1
    INTEGER(C_INT), PARAMETER :: HQ_FP64 = 0, HQ_MIXED = 1
    INTEGER(C_INT), PARAMETER :: HQ_LAYOUT_3N = 0, HQ_LAYOUT_N3 = 1
    ! tuning keys
    INTEGER(C_INT), PARAMETER :: HQ_TUNE_BUNCH = 1, HQ_TUNE_SUPER = 2, HQ_TUNE_FUSED = 3
    INTEGER(C_INT), PARAMETER :: HQ_TUNE_LEAF = 4, HQ_TUNE_FMM_LEAF = 5, HQ_TUNE_GRAPH = 6, &
                                   HQ_TUNE_SPLIT = 7, HQ_TUNE_FAR = 8, HQ_TUNE_FAR_INNER = 9, &
                                   HQ_TUNE_GRAPH_GRID = 10, HQ_TUNE_MAC = 11, &
                                   HQ_TUNE_FMM_BIN_WINDOW = 12
    ! FMM modes
    INTEGER(C_INT), PARAMETER :: HQ_FMM_MFMA = 0, HQ_FMM_FP64 = 1
    ! integration
    INTEGER(C_INT), PARAMETER :: HQ_SCHEME_VERLET = 0, HQ_SCHEME_EULER = 1
    INTEGER(C_INT), PARAMETER :: HQ_FORCE_TREE = 0, HQ_FORCE_BUNCHED = 1, HQ_FORCE_DIRECT = 2
    INTEGER(C_INT), PARAMETER :: HQ_CORR_REFERENCE = 0, HQ_CORR_EXACT = 1

    LOGICAL :: hq_stop_on_error = .TRUE.
    INTEGER(C_INT) :: hq_status = HQ_OK

    ! hq_parareal_params (field order and types as in include/hq.h)
    TYPE, BIND(C) :: hq_parareal_params
        INTEGER(C_INT64_T) :: n = 0
        INTEGER(C_INT) :: num_windows = 0, itermax = 0
        REAL(C_DOUBLE) :: tf = 0, maxerror = 0, G = 1, eps2 = 0
        INTEGER(C_INT) :: coarse_force = 0, coarse_method = 11, coarse_precision = 0
        INTEGER(C_INT) :: coarse_steps = 1
        REAL(C_DOUBLE) :: coarse_theta2 = 0
        INTEGER(C_INT) :: fine_force = 2, fine_method = 10, fine_precision = 0, fine_steps = 1
        REAL(C_DOUBLE) :: fine_theta2 = 0
        INTEGER(C_INT) :: group_size = 64
        INTEGER(C_INT) :: correction = 0
        INTEGER(C_INT) :: coarse_scheme = 1
    END TYPE

    ! hq_tree_info_t
    TYPE, BIND(C) :: hq_tree_info_t
        INTEGER(C_INT64_T) :: n, nnodes, ngroups, lost
        INTEGER(C_INT) :: most_depth, multipole
        INTEGER(C_INT64_T) :: list_cells, list_parts, n_p2p, n_m2p
        REAL(C_DOUBLE) :: ms_build, ms_keys, ms_sort, ms_nodes, ms_upward
        REAL(C_DOUBLE) :: ms_walk, ms_lists, ms_eval
        INTEGER(C_INT64_T) :: exec_cells, exec_parts, max_cells, max_parts, max_stack
        INTEGER(C_INT64_T) :: split_walks, far_cells, walk_stack
    END TYPE

    INTERFACE
        ! ---- f2py `api.api_mod` surface (fortran/api3d.f90) ------------------
        ! api3d.f90:43-72 compute_forces_direct -> F(N,3)
        FUNCTION hq_compute_forces_direct(n, m, x, y, z, G, eps, F) &
                BIND(C, name="hq_compute_forces_direct") RESULT(st)
            IMPORT :: C_INT, C_INT64_T, C_DOUBLE
            INTEGER(C_INT64_T), VALUE :: n
            REAL(C_DOUBLE), INTENT(IN) :: m(*), x(*), y(*), z(*)
            REAL(C_DOUBLE), VALUE :: G, eps
            REAL(C_DOUBLE), INTENT(OUT) :: F(*)
            INTEGER(C_INT) :: st
        END FUNCTION

        ! api3d.f90:20-41 compute_forces -> F(N,3)
        FUNCTION hq_compute_forces(n, m, x, y, z, theta, G, eps, method, legacy_theta2, F) &
                BIND(C, name="hq_compute_forces") RESULT(st)
            IMPORT :: C_INT, C_INT64_T, C_DOUBLE
            INTEGER(C_INT64_T), VALUE :: n
            REAL(C_DOUBLE), INTENT(IN) :: m(*), x(*), y(*), z(*)
            REAL(C_DOUBLE), VALUE :: theta, G, eps
            INTEGER(C_INT), VALUE :: method, legacy_theta2
            REAL(C_DOUBLE), INTENT(OUT) :: F(*)
            INTEGER(C_INT) :: st
        END FUNCTION

        ! api3d.f90:9-18 generate_tree
        FUNCTION hq_generate_tree(n, m, x, y, z, method) BIND(C, name="hq_generate_tree") RESULT(st)
            IMPORT :: C_INT, C_INT64_T, C_DOUBLE
            INTEGER(C_INT64_T), VALUE :: n
            REAL(C_DOUBLE), INTENT(IN) :: m(*), x(*), y(*), z(*)
            INTEGER(C_INT), VALUE :: method
            INTEGER(C_INT) :: st
        END FUNCTION

        ! ---- OctreeType handle (octree_default.f90) ---------------------------
        ! pre_init (:178-282)
        FUNCTION hq_tree_create(h, n, method, group_size, precision) &
                BIND(C, name="hq_tree_create") RESULT(st)
            IMPORT :: C_INT, C_INT64_T, C_PTR
            TYPE(C_PTR), INTENT(OUT) :: h
            INTEGER(C_INT64_T), VALUE :: n
            INTEGER(C_INT), VALUE :: method, group_size, precision
            INTEGER(C_INT) :: st
        END FUNCTION

        ! one OctreeType per system / time window (parareal.f90:28-140)
        FUNCTION hq_tree_create_batch(h, nsys, counts, method, group_size, precision) &
                BIND(C, name="hq_tree_create_batch") RESULT(st)
            IMPORT :: C_INT, C_INT64_T, C_PTR
            TYPE(C_PTR), INTENT(OUT) :: h
            INTEGER(C_INT), VALUE :: nsys
            INTEGER(C_INT64_T), INTENT(IN) :: counts(*)
            INTEGER(C_INT), VALUE :: method, group_size, precision
            INTEGER(C_INT) :: st
        END FUNCTION

        FUNCTION hq_tree_set_masses(h, m, where) BIND(C, name="hq_tree_set_masses") RESULT(st)
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: h
            REAL(C_DOUBLE), INTENT(IN) :: m(*)
            INTEGER(C_INT), VALUE :: where
            INTEGER(C_INT) :: st
        END FUNCTION

        ! init (:320-387)
        FUNCTION hq_tree_build(h, x, y, z, where) BIND(C, name="hq_tree_build") RESULT(st)
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: h
            REAL(C_DOUBLE), INTENT(IN) :: x(*), y(*), z(*)
            INTEGER(C_INT), VALUE :: where
            INTEGER(C_INT) :: st
        END FUNCTION

        ! bh_forces_over_p (:895) for every p, dehnen_eval (:1680), forces_groups (:1811)
        FUNCTION hq_tree_forces(h, theta2, G, eps2, F, layout, where) &
                BIND(C, name="hq_tree_forces") RESULT(st)
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: h
            REAL(C_DOUBLE), VALUE :: theta2, G, eps2
            REAL(C_DOUBLE), INTENT(OUT) :: F(*)
            INTEGER(C_INT), VALUE :: layout, where
            INTEGER(C_INT) :: st
        END FUNCTION

        FUNCTION hq_tree_forces_bunched(h, theta2, G, eps2, F, layout, where, precision) &
                BIND(C, name="hq_tree_forces_bunched") RESULT(st)
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: h
            REAL(C_DOUBLE), VALUE :: theta2, G, eps2
            REAL(C_DOUBLE), INTENT(OUT) :: F(*)
            INTEGER(C_INT), VALUE :: layout, where, precision
            INTEGER(C_INT) :: st
        END FUNCTION

        FUNCTION hq_tree_destroy(h) BIND(C, name="hq_tree_destroy") RESULT(st)
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: h
            INTEGER(C_INT) :: st
        END FUNCTION

        FUNCTION hq_tree_tune(h, key, value) BIND(C, name="hq_tree_tune") RESULT(st)
            IMPORT :: C_INT, C_INT64_T, C_PTR
            TYPE(C_PTR), VALUE :: h
            INTEGER(C_INT), VALUE :: key
            INTEGER(C_INT64_T), VALUE :: value
            INTEGER(C_INT) :: st
        END FUNCTION

        FUNCTION hq_tree_fmm(h, order, theta2, G, eps2, mode, F, layout, where) &
                BIND(C, name="hq_tree_fmm") RESULT(st)
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: h
            INTEGER(C_INT), VALUE :: order
            REAL(C_DOUBLE), VALUE :: theta2, G, eps2
            INTEGER(C_INT), VALUE :: mode
            REAL(C_DOUBLE), INTENT(OUT) :: F(*)
            INTEGER(C_INT), VALUE :: layout, where
            INTEGER(C_INT) :: st
        END FUNCTION

        FUNCTION hq_tree_autotune(h, theta2, G, eps2, precision, reps, bunch, super_size, leaf, ms) &
                BIND(C, name="hq_tree_autotune") RESULT(st)
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: h
            REAL(C_DOUBLE), VALUE :: theta2, G, eps2
            INTEGER(C_INT), VALUE :: precision, reps
            INTEGER(C_INT), INTENT(OUT) :: bunch, super_size, leaf
            REAL(C_DOUBLE), INTENT(OUT) :: ms
            INTEGER(C_INT) :: st
        END FUNCTION

        FUNCTION hq_tree_get_tune(h, key, value) BIND(C, name="hq_tree_get_tune") RESULT(st)
            IMPORT :: C_INT, C_INT64_T, C_PTR
            TYPE(C_PTR), VALUE :: h
            INTEGER(C_INT), VALUE :: key
            INTEGER(C_INT64_T), INTENT(OUT) :: value
            INTEGER(C_INT) :: st
        END FUNCTION

        ! the tuning table; path: a NUL-terminated file name (TRIM(p)//C_NULL_CHAR)
        FUNCTION hq_tune_table_save(path) BIND(C, name="hq_tune_table_save") RESULT(st)
            IMPORT :: C_INT, C_CHAR
            CHARACTER(KIND=C_CHAR), INTENT(IN) :: path(*)
            INTEGER(C_INT) :: st
        END FUNCTION

        FUNCTION hq_tune_table_load(path) BIND(C, name="hq_tune_table_load") RESULT(st)
            IMPORT :: C_INT, C_CHAR
            CHARACTER(KIND=C_CHAR), INTENT(IN) :: path(*)
            INTEGER(C_INT) :: st
        END FUNCTION

        FUNCTION hq_tune_table_clear() BIND(C, name="hq_tune_table_clear") RESULT(st)
            IMPORT :: C_INT
            INTEGER(C_INT) :: st
        END FUNCTION

        ! ---- time integration around the force path ---------------------------
        FUNCTION hq_tree_set_state(h, x, y, z, px, py, pz, where) &
                BIND(C, name="hq_tree_set_state") RESULT(st)
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: h
            REAL(C_DOUBLE), INTENT(IN) :: x(*), y(*), z(*), px(*), py(*), pz(*)
            INTEGER(C_INT), VALUE :: where
            INTEGER(C_INT) :: st
        END FUNCTION

        FUNCTION hq_tree_get_state(h, x, y, z, px, py, pz, where) &
                BIND(C, name="hq_tree_get_state") RESULT(st)
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: h
            REAL(C_DOUBLE), INTENT(OUT) :: x(*), y(*), z(*), px(*), py(*), pz(*)
            INTEGER(C_INT), VALUE :: where
            INTEGER(C_INT) :: st
        END FUNCTION

        ! integration.f90:110-172 (Verlet), parareal.f90:8-44 (Euler)
        FUNCTION hq_tree_integrate(h, scheme, dt, nsteps, theta2, G, eps2, force) &
                BIND(C, name="hq_tree_integrate") RESULT(st)
            IMPORT :: C_INT, C_INT64_T, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: h
            INTEGER(C_INT), VALUE :: scheme
            REAL(C_DOUBLE), VALUE :: dt
            INTEGER(C_INT64_T), VALUE :: nsteps
            REAL(C_DOUBLE), VALUE :: theta2, G, eps2
            INTEGER(C_INT), VALUE :: force
            INTEGER(C_INT) :: st
        END FUNCTION

        ! out(7*nsys): E, P(3), J(3) per system
        FUNCTION hq_tree_energy(h, G, eps2, out) BIND(C, name="hq_tree_energy") RESULT(st)
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: h
            REAL(C_DOUBLE), VALUE :: G, eps2
            REAL(C_DOUBLE), INTENT(OUT) :: out(*)
            INTEGER(C_INT) :: st
        END FUNCTION

        ! ---- harness error metric (benchmarks/test_forces.f90:66-127) ----------
        FUNCTION hq_errors_ma(n, m, x, y, z, F, Fd, layout, where, errors) &
                BIND(C, name="hq_errors_ma") RESULT(st)
            IMPORT :: C_INT, C_INT64_T, C_DOUBLE
            INTEGER(C_INT64_T), VALUE :: n
            REAL(C_DOUBLE), INTENT(IN) :: m(*), x(*), y(*), z(*), F(*), Fd(*)
            INTEGER(C_INT), VALUE :: layout, where
            REAL(C_DOUBLE), INTENT(OUT) :: errors(7)
            INTEGER(C_INT) :: st
        END FUNCTION

        ! ---- Parareal (parareal/parareal.f90:97-186) ---------------------------
        FUNCTION hq_parareal(prm, m, x0, y0, z0, px0, py0, pz0, where, U_out, out_where, errors, &
                             iterations) BIND(C, name="hq_parareal") RESULT(st)
            IMPORT :: C_INT, C_DOUBLE, hq_parareal_params
            TYPE(hq_parareal_params), INTENT(IN) :: prm
            REAL(C_DOUBLE), INTENT(IN) :: m(*), x0(*), y0(*), z0(*), px0(*), py0(*), pz0(*)
            INTEGER(C_INT), VALUE :: where
            REAL(C_DOUBLE), INTENT(OUT) :: U_out(*)
            INTEGER(C_INT), VALUE :: out_where
            REAL(C_DOUBLE), INTENT(OUT) :: errors(*)
            INTEGER(C_INT), INTENT(OUT) :: iterations
            INTEGER(C_INT) :: st
        END FUNCTION

        ! ---- inspection ---------------------------------------------------------
        FUNCTION hq_tree_info(h, info) BIND(C, name="hq_tree_info") RESULT(st)
            IMPORT :: C_INT, C_PTR, hq_tree_info_t
            TYPE(C_PTR), VALUE :: h
            TYPE(hq_tree_info_t), INTENT(OUT) :: info
            INTEGER(C_INT) :: st
        END FUNCTION

        ! node store in level order; any pointer may be C_NULL_PTR
        FUNCTION hq_tree_export(h, depth, type, particle, child_start, key_hi, key_lo, mass, com, &
                                quad, oct, center) BIND(C, name="hq_tree_export") RESULT(st)
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: h, depth, type, particle, child_start, key_hi, key_lo
            TYPE(C_PTR), VALUE :: mass, com, quad, oct, center
            INTEGER(C_INT) :: st
        END FUNCTION

        ! order(i) = original 0-based index of sorted slot i
        FUNCTION hq_tree_order(h, order) BIND(C, name="hq_tree_order") RESULT(st)
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: h
            INTEGER(C_INT), INTENT(OUT) :: order(*)
            INTEGER(C_INT) :: st
        END FUNCTION

        ! ---- direct sum on device memory (fortran/direct_forces.f90:8-64) ------
        FUNCTION hq_direct_forces_device(n, m, x, y, z, nt, targets, G, eps2, precision, F) &
                BIND(C, name="hq_direct_forces_device") RESULT(st)
            IMPORT :: C_INT, C_INT64_T, C_PTR, C_DOUBLE
            INTEGER(C_INT64_T), VALUE :: n
            TYPE(C_PTR), VALUE :: m, x, y, z
            INTEGER(C_INT64_T), VALUE :: nt
            TYPE(C_PTR), VALUE :: targets
            REAL(C_DOUBLE), VALUE :: G, eps2
            INTEGER(C_INT), VALUE :: precision
            TYPE(C_PTR), VALUE :: F
            INTEGER(C_INT) :: st
        END FUNCTION

        ! ---- Morton tree (fortran/morton/morton_octree.f90) --------------------
        FUNCTION hq_morton_keys(n, m, q, L, ssa, keys, perm, L_out, where) &
                BIND(C, name="hq_morton_keys") RESULT(st)
            IMPORT :: C_INT, C_INT64_T, C_DOUBLE
            INTEGER(C_INT64_T), VALUE :: n
            REAL(C_DOUBLE), INTENT(IN) :: m(*), q(*)
            INTEGER(C_INT), VALUE :: L
            REAL(C_DOUBLE), VALUE :: ssa
            INTEGER(C_INT64_T), INTENT(OUT) :: keys(*)
            INTEGER(C_INT), INTENT(OUT) :: perm(*), L_out
            INTEGER(C_INT), VALUE :: where
            INTEGER(C_INT) :: st
        END FUNCTION

        ! hq_morton_keys on device arrays with no device allocation (no
        ! reference counterpart; for callers that cannot allocate)
        FUNCTION hq_morton_keys_noalloc(n, m, q, L, ssa, keys, perm, L_out) &
                BIND(C, name="hq_morton_keys_noalloc") RESULT(st)
            IMPORT :: C_INT, C_INT64_T, C_PTR, C_DOUBLE
            INTEGER(C_INT64_T), VALUE :: n
            TYPE(C_PTR), VALUE :: m, q
            INTEGER(C_INT), VALUE :: L
            REAL(C_DOUBLE), VALUE :: ssa
            TYPE(C_PTR), VALUE :: keys, perm
            INTEGER(C_INT), INTENT(OUT) :: L_out
            INTEGER(C_INT) :: st
        END FUNCTION

        FUNCTION hq_morton_create(h, n, L, ssa, use_quad) BIND(C, name="hq_morton_create") RESULT(st)
            IMPORT :: C_INT, C_INT64_T, C_PTR, C_DOUBLE
            TYPE(C_PTR), INTENT(OUT) :: h
            INTEGER(C_INT64_T), VALUE :: n
            INTEGER(C_INT), VALUE :: L
            REAL(C_DOUBLE), VALUE :: ssa
            INTEGER(C_INT), VALUE :: use_quad
            INTEGER(C_INT) :: st
        END FUNCTION

        FUNCTION hq_morton_build(h, m, q, where) BIND(C, name="hq_morton_build") RESULT(st)
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: h
            REAL(C_DOUBLE), INTENT(IN) :: m(*), q(*)
            INTEGER(C_INT), VALUE :: where
            INTEGER(C_INT) :: st
        END FUNCTION

        FUNCTION hq_morton_forces(h, G, eps2, theta2, mult, F, layout, where) &
                BIND(C, name="hq_morton_forces") RESULT(st)
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: h
            REAL(C_DOUBLE), VALUE :: G, eps2, theta2
            INTEGER(C_INT), VALUE :: mult
            REAL(C_DOUBLE), INTENT(OUT) :: F(*)
            INTEGER(C_INT), VALUE :: layout, where
            INTEGER(C_INT) :: st
        END FUNCTION

        FUNCTION hq_morton_info(h, L, nnodes) BIND(C, name="hq_morton_info") RESULT(st)
            IMPORT :: C_INT, C_INT64_T, C_PTR
            TYPE(C_PTR), VALUE :: h
            INTEGER(C_INT), INTENT(OUT) :: L
            INTEGER(C_INT64_T), INTENT(OUT) :: nnodes
            INTEGER(C_INT) :: st
        END FUNCTION

        ! any pointer may be C_NULL_PTR
        FUNCTION hq_morton_export(h, perm, prefix, level, first, last, parent, nchild, &
                                  first_child, next, mass, qcm, quad) &
                BIND(C, name="hq_morton_export") RESULT(st)
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: h, perm, prefix, level, first, last, parent, nchild
            TYPE(C_PTR), VALUE :: first_child, next, mass, qcm, quad
            INTEGER(C_INT) :: st
        END FUNCTION

        FUNCTION hq_morton_destroy(h) BIND(C, name="hq_morton_destroy") RESULT(st)
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: h
            INTEGER(C_INT) :: st
        END FUNCTION

        ! ---- device memory plumbing (no reference counterpart) -----------------
        FUNCTION hq_dev_alloc(p, bytes) BIND(C, name="hq_dev_alloc") RESULT(st)
            IMPORT :: C_INT, C_INT64_T, C_PTR
            TYPE(C_PTR), INTENT(OUT) :: p
            INTEGER(C_INT64_T), VALUE :: bytes
            INTEGER(C_INT) :: st
        END FUNCTION

        FUNCTION hq_dev_free(p) BIND(C, name="hq_dev_free") RESULT(st)
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: p
            INTEGER(C_INT) :: st
        END FUNCTION

        ! kind: 0 host->device, 1 device->host, 2 device->device
        FUNCTION hq_dev_copy(dst, src, bytes, kind) BIND(C, name="hq_dev_copy") RESULT(st)
            IMPORT :: C_INT, C_INT64_T, C_PTR
            TYPE(C_PTR), VALUE :: dst, src
            INTEGER(C_INT64_T), VALUE :: bytes
            INTEGER(C_INT), VALUE :: kind
            INTEGER(C_INT) :: st
        END FUNCTION

        FUNCTION hq_dev_sync() BIND(C, name="hq_dev_sync") RESULT(st)
            IMPORT :: C_INT
            INTEGER(C_INT) :: st
        END FUNCTION

        FUNCTION hq_last_error() BIND(C, name="hq_last_error") RESULT(p)
            IMPORT :: C_PTR
            TYPE(C_PTR) :: p
        END FUNCTION

        FUNCTION hq_version() BIND(C, name="hq_version") RESULT(p)
            IMPORT :: C_PTR
            TYPE(C_PTR) :: p
        END FUNCTION

        FUNCTION c_strlen(s) BIND(C, name="strlen") RESULT(k)
            IMPORT :: C_PTR, C_SIZE_T
            TYPE(C_PTR), VALUE :: s
            INTEGER(C_SIZE_T) :: k
        END FUNCTION
    END INTERFACE

CONTAINS

    ! A C string returned by the library as a Fortran string.
    FUNCTION hq_c_string(p) RESULT(s)
        TYPE(C_PTR), INTENT(IN) :: p
        CHARACTER(LEN=:), ALLOCATABLE :: s
        CHARACTER(KIND=C_CHAR), POINTER :: c(:)
        INTEGER :: k, i
        IF (.NOT. C_ASSOCIATED(p)) THEN
            s = ''
            RETURN
        END IF
        k = INT(c_strlen(p))
        CALL C_F_POINTER(p, c, [k])
        ALLOCATE(CHARACTER(LEN=k) :: s)
        DO i = 1, k
            s(i:i) = c(i)
        END DO
    END FUNCTION

    ! The calling thread's last error message.
    FUNCTION hq_error_message() RESULT(s)
        CHARACTER(LEN=:), ALLOCATABLE :: s
        s = hq_c_string(hq_last_error())
    END FUNCTION

    ! Error handling of the Fortran mirrors, whose signatures (the reference's)
    ! carry no status: with hq_stop_on_error (the default) a failing call stops
    ! with the library's message, as the reference's routines STOP on their own
    ! errors (octree_default.f90:457-460); otherwise the status is kept in
    ! hq_status (first failure since the caller last cleared it) and the call
    ! returns with its outputs undefined.
    SUBROUTINE hq_check(st, what)
        INTEGER(C_INT), INTENT(IN) :: st
        CHARACTER(LEN=*), INTENT(IN) :: what
        IF (st == HQ_OK) RETURN
        IF (hq_status == HQ_OK) hq_status = st
        IF (hq_stop_on_error) THEN
            WRITE(*, '(A, A, I0, A, A)') what, ': status ', st, ': ', hq_error_message()
            ERROR STOP 1
        END IF
    END SUBROUTINE
END MODULE
