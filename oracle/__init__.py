"""CPU oracle for the TigerCI sigma-vector path (sigma = H.c inside Davidson-Liu).

TEST INFRASTRUCTURE ONLY.  Nothing in the product path (``tigerci_amd``) imports,
calls or links anything in this package.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use it,
and only as the checker (or as the timed CPU baseline), never as the thing that
is measured or shipped.

Contents
--------
problem   synthetic closed-shell MRSDCI problem: orbital energies, Cholesky
          vectors L^J_pq, local pair domains, the CSF layout of the CI vector
          (graph_mod.f90:1041-1092 layout) and seeded CI vectors.
scepper   pack/unpack of pair blocks (scepmaker_mod.f90:60-798) and a literal
          restatement of the 4-external ladder per (a,c) and walk
          (three_and_four_seg_mod_1.f90:3240-3627).
sigma     dense pair-block restatement of every sigma term of htimesc
          (htimesc.f90:46-147) for the closed-shell single-reference CSF space,
          plus the exact diagonal H_ii (diag_element_mod_2.f90:83 role).
davidson  Davidson-Liu restatement (davidson.f90:47-239, 478-723).
drt       the distinct row table, lexical walk order and the walk-keeping
          tests of the internal tree search (graph_mod.f90:101-380,963-1028,
          new_tree_search.f90:45-202).
spin      SGGA spin tables and contracted cycles (spin_mod.f90:99-1312).
domains   hole pairs, allowed virtuals and the weak-pairs test
          (locist_mod.f90:210-245,453-607, allowed_virtuals_mod.f90:163-219).
cholesky  the AO -> MO transformation of Cholesky vectors
          (cholesky_transform.f90:305-411).
records   the ladder loop records of the iteration-1 dry run
          (three_and_four_seg_mod_1.f90:318-469,3142-3238).
cpu_ladder / cpu_sigma  the reference ladder loop in C with OpenMP and BLAS
          (the timed CPU baseline) and its Python driver.

How each function is pinned (the checks live in ``tests/``):
  * every sigma term and the diagonal: brute-force Slater-determinant CI with
    Slater-Condon rules on tiny spaces, projected onto explicitly constructed
    spin-adapted CSFs (tests/detci.py) -- independent of the closed forms here;
  * the CSF basis: orthonormality and S^2 = 0 in determinant space;
  * pack/unpack: round-trip identities and adjointness;
  * packed ladder vs dense ladder: agreement to round-off;
  * Davidson: converged roots equal numpy.linalg.eigh of the dense H;
  * hermiticity c1.(H c2) == c2.(H c1) and H_ii == (H e_i)_i.
  * the DRT, walk order and kept-walk counts: brute-force walk enumeration
    and the counts the reference's test outputs print (Methane SDCI and LSDCI,
    methanol MRACPF);
  * the AO -> MO transform: an independent dense transformation and the
    four-index identity (pq|rs) = sum_J L_pq L_rs against transformed AO integrals;
  * allowed virtuals: the methane LSDCI truncation counts and an independent
    statement by hole/particle counting on seeded multireference spaces;
  * ladder records: the closed-form count (every unordered pair of a walk's
    virtuals), positions that point back at the record's virtuals, the
    reference's order, and ignorable pairs removing exactly their records.
Parity unpinned: the phase of the i = j pair walks relative to the valence
CSF (the loop types that would fix it are not restated).  The CSF phase of
i > j pair walks is pinned by the restated valence <-> N-2 coupling
coefficients, and that of the N-1 walks by the valence <-> N-1 coefficient
(+sqrt 2 times the reference's G(a), tests/test_spin.py).
"""
