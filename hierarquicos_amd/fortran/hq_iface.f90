! hq_iface.f90 -- Fortran interface to libhq.so (include/hq.h) through
! ISO_C_BINDING, for Fortran drivers of the reference (benchmarks/test_forces.f90,
! parareal/parareal.f90) that call compute_forces / compute_forces_direct /
! OctreeType.  Link the driver with -L<dir of libhq.so> -lhq.
MODULE hq_iface
    USE ISO_C_BINDING
    IMPLICIT NONE
    INTEGER(C_INT), PARAMETER :: HQ_OK = 0, HQ_HOST = 0, HQ_DEVICE = 1
    INTEGER(C_INT), PARAMETER :: HQ_FP64 = 0, HQ_MIXED = 1
    INTEGER(C_INT), PARAMETER :: HQ_LAYOUT_3N = 0, HQ_LAYOUT_N3 = 1

    INTERFACE
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

        ! OctreeType: pre_init / init / forces over all p (3,N)
        FUNCTION hq_tree_create(h, n, method, group_size, precision) &
                BIND(C, name="hq_tree_create") RESULT(st)
            IMPORT :: C_INT, C_INT64_T, C_PTR
            TYPE(C_PTR), INTENT(OUT) :: h
            INTEGER(C_INT64_T), VALUE :: n
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

        FUNCTION hq_tree_build(h, x, y, z, where) BIND(C, name="hq_tree_build") RESULT(st)
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: h
            REAL(C_DOUBLE), INTENT(IN) :: x(*), y(*), z(*)
            INTEGER(C_INT), VALUE :: where
            INTEGER(C_INT) :: st
        END FUNCTION

        FUNCTION hq_tree_forces(h, theta2, G, eps2, F, layout, where) &
                BIND(C, name="hq_tree_forces") RESULT(st)
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: h
            REAL(C_DOUBLE), VALUE :: theta2, G, eps2
            REAL(C_DOUBLE), INTENT(OUT) :: F(*)
            INTEGER(C_INT), VALUE :: layout, where
            INTEGER(C_INT) :: st
        END FUNCTION

        FUNCTION hq_tree_destroy(h) BIND(C, name="hq_tree_destroy") RESULT(st)
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: h
            INTEGER(C_INT) :: st
        END FUNCTION

        FUNCTION hq_last_error() BIND(C, name="hq_last_error") RESULT(p)
            IMPORT :: C_PTR
            TYPE(C_PTR) :: p
        END FUNCTION
    END INTERFACE
END MODULE
