! ref_shim.f90 -- C bindings over the reference's own Fortran modules, used only
! to validate the C restatement (hq_oracle.c) and to write golden fixtures.
! TEST INFRASTRUCTURE ONLY.  It is compiled together with the reference
! sources where they lie (fortran/octree_default.f90, fortran/direct_forces.f90,
! fortran/morton/morton_octree.f90); see oracle/Makefile.  Output goes to
! oracle/_ref/ and never into the product.
MODULE hq_ref_shim
    USE ISO_C_BINDING
    USE octree_mod
    USE direct_forces_mod
    USE morton_octree_mod
    IMPLICIT NONE
    TYPE(OctreeType), ALLOCATABLE, SAVE :: tree
    TYPE(morton_tree_type), ALLOCATABLE, SAVE :: mtree
CONTAINS

! direct_forces.f90:8 with nt = 1 (serial pair loop); F is (3,n)
SUBROUTINE ref_direct(n, m, x, y, z, G, eps2, F) BIND(C, name="ref_direct")
    INTEGER(C_INT), VALUE :: n
    REAL(C_DOUBLE), INTENT(IN) :: m(n), x(n), y(n), z(n)
    REAL(C_DOUBLE), VALUE :: G, eps2
    REAL(C_DOUBLE), INTENT(OUT) :: F(3, n)
    F = compute_forces_direct(m, x, y, z, G, eps2, 1)
END SUBROUTINE

! pre_init + init of OctreeType (octree_default.f90:178, :320); returns nnodes
FUNCTION ref_tree_build(n, m, x, y, z, method, group_size) RESULT(nn) BIND(C, name="ref_tree_build")
    INTEGER(C_INT), VALUE :: n, method, group_size
    REAL(C_DOUBLE), INTENT(IN) :: m(n), x(n), y(n), z(n)
    INTEGER(C_INT) :: nn
    IF (ALLOCATED(tree)) DEALLOCATE(tree)
    ALLOCATE(tree)
    CALL tree % pre_init(m, method, group_size=group_size)
    CALL tree % init(x, y, z)
    nn = tree % number_of_nodes
END FUNCTION

SUBROUTINE ref_tree_dump(cx, cy, cz, half, L2, mass, qx, qy, qz, particle, ntype, depth, child, &
                         quad, oct) BIND(C, name="ref_tree_dump")
    REAL(C_DOUBLE), INTENT(OUT) :: cx(*), cy(*), cz(*), half(*), L2(*), mass(*), qx(*), qy(*), qz(*)
    INTEGER(C_INT), INTENT(OUT) :: particle(*), ntype(*), depth(*), child(8, *)
    REAL(C_DOUBLE), INTENT(OUT) :: quad(6, *), oct(14, *)
    INTEGER :: nn, k
    nn = tree % number_of_nodes
    cx(1:nn) = tree % ns_cx(1:nn)
    cy(1:nn) = tree % ns_cy(1:nn)
    cz(1:nn) = tree % ns_cz(1:nn)
    half(1:nn) = tree % ns_halfside(1:nn)
    L2(1:nn) = tree % ns_L2(1:nn)
    mass(1:nn) = tree % ns_mass(1:nn)
    qx(1:nn) = tree % ns_qcm_x(1:nn)
    qy(1:nn) = tree % ns_qcm_y(1:nn)
    qz(1:nn) = tree % ns_qcm_z(1:nn)
    ntype(1:nn) = tree % ns_type(1:nn)
    depth(1:nn) = tree % ns_depth(1:nn)
    DO k = 1, nn
        ! 0-based particle and child indices for C
        particle(k) = MERGE(tree % ns_particle(k) - 1, -1, tree % ns_particle(k) > 0)
        child(:, k) = MERGE(tree % ns_child(:, k) - 1, -1, tree % ns_child(:, k) > 0)
    END DO
    IF (tree % multipole >= 4) quad(:, 1:nn) = tree % ns_quad(:, 1:nn)
    IF (tree % multipole >= 8) oct(:, 1:nn) = tree % ns_oct(:, 1:nn)
END SUBROUTINE

! bh_forces_over_p (octree_default.f90:895) for every particle; F is (3,n)
SUBROUTINE ref_tree_bh(theta2, G, eps2, F) BIND(C, name="ref_tree_bh")
    REAL(C_DOUBLE), VALUE :: theta2, G, eps2
    REAL(C_DOUBLE), INTENT(OUT) :: F(3, *)
    INTEGER :: p
    !$OMP PARALLEL DO SCHEDULE(DYNAMIC, 64) SHARED(tree, F, theta2, G, eps2) PRIVATE(p)
    DO p = 1, tree % N
        F(:, p) = tree % forces(p, theta2, G, eps2)
    END DO
    !$OMP END PARALLEL DO
END SUBROUTINE

! bh_forces_over_p for a subset of targets (0-based indices); F is (3,nt)
SUBROUTINE ref_tree_bh_targets(nt, targets, theta2, G, eps2, F) BIND(C, name="ref_tree_bh_targets")
    INTEGER(C_INT), VALUE :: nt
    INTEGER(C_INT64_T), INTENT(IN) :: targets(nt)
    REAL(C_DOUBLE), VALUE :: theta2, G, eps2
    REAL(C_DOUBLE), INTENT(OUT) :: F(3, nt)
    INTEGER :: i
    !$OMP PARALLEL DO SCHEDULE(DYNAMIC, 64) SHARED(tree, F, targets, theta2, G, eps2) PRIVATE(i)
    DO i = 1, nt
        F(:, i) = tree % forces(INT(targets(i)) + 1, theta2, G, eps2)
    END DO
    !$OMP END PARALLEL DO
END SUBROUTINE

! forces_groups (octree_default.f90:1811); F is (3,n)
SUBROUTINE ref_tree_groups(theta2, G, eps2, F, ngroups) BIND(C, name="ref_tree_groups")
    REAL(C_DOUBLE), VALUE :: theta2, G, eps2
    REAL(C_DOUBLE), INTENT(OUT) :: F(3, tree % N)
    INTEGER(C_INT), INTENT(OUT) :: ngroups
    CALL tree % forces_groups(theta2, G, eps2, F, 1)
    ngroups = tree % ngroups
END SUBROUTINE

! dehnen_eval (octree_default.f90:1680) with the default thread count; F is (3,n)
SUBROUTINE ref_tree_dehnen(theta2, eps2, G, F) BIND(C, name="ref_tree_dehnen")
    REAL(C_DOUBLE), VALUE :: theta2, eps2, G
    REAL(C_DOUBLE), INTENT(OUT) :: F(3, tree % N)
    CALL tree % dehnen_eval(theta2, eps2, G, F)
END SUBROUTINE

SUBROUTINE ref_tree_free() BIND(C, name="ref_tree_free")
    IF (ALLOCATED(tree)) DEALLOCATE(tree)
END SUBROUTINE

! morton_tree_type%init (morton_octree.f90:73); returns the depth used
FUNCTION ref_morton_build(n, m, q, L, ssa, use_quad, nnodes) RESULT(Lout) BIND(C, name="ref_morton_build")
    INTEGER(C_INT), VALUE :: n, L, use_quad
    REAL(C_DOUBLE), INTENT(IN) :: m(n), q(3, n)
    REAL(C_DOUBLE), VALUE :: ssa
    INTEGER(C_INT), INTENT(OUT) :: nnodes
    INTEGER(C_INT) :: Lout
    IF (ALLOCATED(mtree)) DEALLOCATE(mtree)
    ALLOCATE(mtree)
    CALL mtree % init(n, m, q, L, ssa, use_quad /= 0)
    Lout = mtree % depth
    nnodes = mtree % nnodes
END FUNCTION

! permutation (0-based) and node records of the Morton tree
SUBROUTINE ref_morton_dump(perm, prefix, level, first, last, parent, nchild, child, next, &
                           size2, mass, qcm, quad) BIND(C, name="ref_morton_dump")
    INTEGER(C_INT), INTENT(OUT) :: perm(*), level(*), first(*), last(*), parent(*), nchild(*)
    INTEGER(C_INT), INTENT(OUT) :: child(8, *), next(*)
    INTEGER(C_INT64_T), INTENT(OUT) :: prefix(*)
    REAL(C_DOUBLE), INTENT(OUT) :: size2(*), mass(*), qcm(3, *), quad(6, *)
    INTEGER :: k, n
    n = mtree % N
    perm(1:n) = mtree % perm(1:n) - 1
    DO k = 1, mtree % nnodes
        prefix(k) = mtree % nodes(k) % prefix
        level(k) = mtree % nodes(k) % level
        first(k) = mtree % nodes(k) % first - 1
        last(k) = mtree % nodes(k) % last - 1
        parent(k) = MERGE(mtree % nodes(k) % parent - 1, -1, mtree % nodes(k) % parent > 0)
        nchild(k) = mtree % nodes(k) % nchild
        child(:, k) = MERGE(mtree % nodes(k) % child(:) - 1, -1, mtree % nodes(k) % child(:) > 0)
        next(k) = mtree % nodes(k) % next - 1
        size2(k) = mtree % nodes(k) % size2
        mass(k) = mtree % nodes(k) % mass
        qcm(:, k) = mtree % nodes(k) % qcm
        quad(:, k) = mtree % nodes(k) % quad
    END DO
END SUBROUTINE

! forces_seq (morton_octree.f90:131); F is (3,n) in original order
SUBROUTINE ref_morton_forces(G, eps2, theta2, mult, F) BIND(C, name="ref_morton_forces")
    REAL(C_DOUBLE), VALUE :: G, eps2, theta2
    INTEGER(C_INT), VALUE :: mult
    REAL(C_DOUBLE), INTENT(OUT) :: F(3, mtree % N)
    F = mtree % forces_seq(G, eps2, theta2, mult /= 0)
END SUBROUTINE

SUBROUTINE ref_morton_free() BIND(C, name="ref_morton_free")
    IF (ALLOCATED(mtree)) DEALLOCATE(mtree)
END SUBROUTINE

END MODULE
