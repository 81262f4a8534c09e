"""CPU oracle for the halo-finding hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg
may import this package, and only as the checker (or the timed CPU baseline).  The
product package ``pyhalma_amd`` never imports it; the product path fails loudly when
its HIP library is missing instead of falling back to anything here.

Each function restates the reference's algorithm and cites the reference lines it
follows.  How each one is pinned (golden vectors, closed forms, brute force, the
reference's own Fortran built under ``oracle/_ref``) is recorded in the tests that
exercise it; a function without such a check would say "parity unpinned" here.
"""
from .native import (  # noqa: F401
    lib_path,
    pot_sum_f32,
    pot_sum_f64,
    pot_soft_f64,
    acc_soft_f64,
    fof_labels_c,
    fof_labels_cells,
    ref_pot_sum_f32,
    have_ref,
)
from .fof import (  # noqa: F401
    fof_bfs,
    fof_labels_scipy,
    fof_labels_brute,
    labels_to_groups,
    fof_parallel,
)
from .reductions import (  # noqa: F401
    total_mass,
    center_of_mass,
    cm_velocity,
    furthest_particle,
    density_peak,
    linspace_ref,
)
from .cleaning import phase_space_clean  # noqa: F401
from .progenitors import main_progenitors, progenitors  # noqa: F401
from .substructure import (  # noqa: F401
    half_mass_radius,
    split_substructure,
    sub_fof,
)
from .unbind import (  # noqa: F401
    G_CONST,
    binding_energy,
    energy_test,
    unbind_iterative,
    rps,
    most_bound_particle,
)
