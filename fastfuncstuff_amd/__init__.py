"""MI355X-native ARMA(1,1) REML engine with the fastfuncstuff ``fit_glm_arma11`` API."""

from .glm.arma import (  # noqa: F401
    ARMA11Results,
    RemlGrid,
    build_censor_run_info,
    compute_arma_lambda,
    ensure_zero_in_grid,
    estimate_valid_grid_pairs,
    fit_glm_arma11,
    get_default_arma_grids,
    precompute_reml_grid,
    search_voxels_precomputed_grid,
    select_arma_params_hierarchical_from_surface,
)

__version__ = "0.1.0"
