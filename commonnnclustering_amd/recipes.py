"""Component registry and builder (cnnclustering recipes.py).

Same key grammar as the reference: dotted component paths
(``fitter.ngetter.dgetter.metric``), ``__`` accepted for ``.``, the short
aliases of COMPONENT_ALT_KW_MAP (recipes.py:366-392), and recursive
construction through ``get_builder_kwargs`` (recipes.py:211-262).  The
reference's type names (``bfs``, ``brute_force``, ``vuset``, ``switch``,
``fifo``, ``euclidean_r`` ...) resolve to the GPU components, so reference
recipes build GPU fitters unchanged.
"""

from __future__ import annotations

from collections.abc import Iterable

import numpy as np

from ._primitive_types import P_AINDEX, P_AVALUE
from . import _types, _fit


def prepare_pass(data):
    """Dummy preparation hook (recipes.py:36-49)."""
    return (data,), {}


def prepare_points_from_parts(data):
    """Normalise point input to one (n, d) float64 array (recipes.py:52-147).

    Accepts a single point (sequence of d numbers), a 2-D sequence of n points,
    or a sequence of 2-D parts of equal width (``meta["edges"]`` records the
    part lengths).  A torch CUDA tensor passes through unchanged (device data
    stays on the device).
    """
    try:
        import torch

        if isinstance(data, torch.Tensor):
            if data.dim() != 2:
                raise ValueError("Dimension mismatch")
            return (data,), {"meta": {"edges": [int(data.shape[0])]}}
    except ImportError:  # pragma: no cover
        pass

    d1 = len(data)
    finished = False
    all_d2_equal = all_d3_equal = False
    if d1 == 0:
        data = [np.array([[]])]
        finished = True
    if not finished:
        try:
            d2 = [len(x) for x in data]
            all_d2_equal = len(set(d2)) == 1
        except TypeError:
            data = [np.array([data])]
            finished = True
    if not finished:
        try:
            d3 = [len(y) for x in data for y in x]
            all_d3_equal = len(set(d3)) == 1
        except TypeError:
            if not all_d2_equal:
                raise ValueError("Dimension mismatch")
            data = [np.asarray(data)]
            finished = True
    if not finished:
        if not all_d3_equal:
            raise ValueError("Dimension mismatch")
        data = [np.asarray(x) for x in data]
    meta = {"edges": [x.shape[0] for x in data]}
    return (np.asarray(np.vstack(data), order="C", dtype=P_AVALUE),), {"meta": meta}


def prepare_neighbourhoods(data):
    """Pad neighbour lists to an (n, max_deg) index array plus counts (recipes.py:150-179)."""
    n_neighbours = [len(s) for s in data]
    pad_to = max(n_neighbours) if n_neighbours else 0
    padded = [np.pad(np.asarray(a, dtype=P_AINDEX), (0, pad_to - n_neighbours[i]), mode="constant",
                     constant_values=0) for i, a in enumerate(data)]
    arr = np.asarray(padded, order="C", dtype=P_AINDEX).reshape(len(n_neighbours), pad_to)
    return (arr, np.asarray(n_neighbours, dtype=P_AINDEX)), {"meta": {}}


class Recipe(dict):
    def __init__(self, valid_input=None):
        super().__init__()
        self.valid_input = [] if valid_input is None else valid_input

    @property
    def valid_input(self):
        return self._valid_input

    @valid_input.setter
    def valid_input(self, values):
        if not isinstance(values, Iterable):
            values = [values]
        checked = []
        for v in values:
            if not (isinstance(v, type) and issubclass(v, _types.InputData)):
                try:
                    v = COMPONENT_NAME_TYPE_MAP["input_data"][v]
                except KeyError:
                    raise ValueError(f"Object {v!r} is not a registered InputData type") from None
            checked.append(v)
        self._valid_input = checked


class Builder:
    """Builds components from a recipe (recipes.py:182-313)."""

    default_recipe = "coordinates"
    default_preparation_hook = staticmethod(prepare_pass)

    def __init__(self, recipe=None, **kwargs):
        if recipe is None:
            recipe = self.default_recipe
        if isinstance(recipe, str):
            recipe = get_registered_recipe(recipe)
        recipe = dict(recipe)
        recipe.update(kwargs)
        recipe = {k.replace("__", "."): v for k, v in recipe.items()}
        self.recipe = {".".join(COMPONENT_ALT_KW_MAP.get(kw, kw) for kw in k.split(".")): v
                       for k, v in recipe.items()}

    def make_component(self, component_kw, alternative=None, prev_kw=None):
        if prev_kw is None:
            prev_kw = []
        full_kw = ".".join([*prev_kw, component_kw])
        if alternative is not None:
            alternative = full_kw.rsplit(".", 1)[0] + f".{alternative}"
        if full_kw in self.recipe:
            details = self.recipe[full_kw]
        elif alternative is not None and alternative in self.recipe:
            details = self.recipe[alternative]
        else:
            return object
        if details is None:
            return None
        args, kwargs = (), {}
        type_kw = COMPONENT_KW_TYPE_ALIAS_MAP.get(component_kw, component_kw)
        if isinstance(details, str):
            component_type = _lookup(type_kw, details)
        elif isinstance(details, tuple):
            component_type, args, kwargs = details
            kwargs = dict(kwargs)
            if isinstance(component_type, str):
                component_type = _lookup(type_kw, component_type)
        elif isinstance(details, type) or callable(details):
            component_type = details
        else:
            return details  # an already constructed component
        if hasattr(component_type, "get_builder_kwargs"):
            for sub_kw, alt in component_type.get_builder_kwargs():
                component = self.make_component(sub_kw, alt, prev_kw=full_kw.split("."))
                if component is not object:
                    kwargs[sub_kw] = component
        return component_type(*args, **kwargs)

    def make_input_data(self, data, preparation_hook=None, details=None):
        if preparation_hook is None:
            preparation_hook = self.recipe.get("preparation_hook", self.default_preparation_hook)
            if isinstance(preparation_hook, str):
                preparation_hook = _lookup("preparation_hook", preparation_hook)
        if details is None:
            try:
                details = self.recipe["input_data"]
            except KeyError:
                raise LookupError("Input data type details not found") from None
        args, kwargs = (), {}
        if isinstance(details, str):
            component_type = _lookup("input_data", details)
        elif isinstance(details, tuple):
            component_type, args, kwargs = details
            kwargs = dict(kwargs)
            if isinstance(component_type, str):
                component_type = _lookup("input_data", component_type)
        else:
            component_type = details
        data_args, data_kwargs = preparation_hook(data)
        args = (*data_args, *args)
        data_kwargs.setdefault("meta", {})
        data_kwargs["meta"].update(kwargs.get("meta", {}))
        kwargs.update(data_kwargs)
        return component_type(*args, **kwargs)


def _lookup(kind, name):
    try:
        return COMPONENT_NAME_TYPE_MAP[kind][name]
    except KeyError:
        if name in UNSUPPORTED_COMPONENTS.get(kind, {}):
            raise NotImplementedError(UNSUPPORTED_COMPONENTS[kind][name]) from None
        raise KeyError(f"no registered {kind} named {name!r}") from None


REGISTERED_RECIPES = {
    "none": {},
    "coordinates": {
        "input_data": "components_mview",
        "preparation_hook": "points_from_parts",
        "fitter": "bfs",
        "fitter.ngetter": "brute_force",
        "fitter.na": "vuset",
        "fitter.checker": "switch",
        "fitter.queue": "fifo",
        "fitter.ngetter.dgetter": "metric",
        "fitter.ngetter.dgetter.metric": "euclidean_r",
        "predictor": "firstmatch",
        "predictor.ngetter": "brute_force",
        "predictor.checker": "switch",
    },
    # FP32 distances with the FP64 guard-band recheck: identical neighbour sets
    "coordinates_f32": {
        "input_data": "components_mview",
        "preparation_hook": "points_from_parts",
        "fitter": "bfs",
        "fitter.ngetter": ("brute_force", (), {"precision": "f32_band"}),
        "fitter.na": "vuset",
        "fitter.checker": "switch",
        "fitter.queue": "fifo",
        "fitter.ngetter.dgetter": "metric",
        "fitter.ngetter.dgetter.metric": "euclidean_r",
        "predictor": "firstmatch",
        "predictor.ngetter": ("brute_force", (), {"precision": "f32_band"}),
        "predictor.checker": "switch",
    },
    "distances": {
        "input_data": "components_mview",
        "preparation_hook": "points_from_parts",
        "fitter": "bfs",
        "fitter.ngetter": "brute_force",
        "fitter.na": "vuset",
        "fitter.checker": "switch",
        "fitter.queue": "fifo",
        "fitter.ngetter.dgetter": "metric",
        "fitter.ngetter.dgetter.metric": "precomputed",
    },
    "neighbourhoods": {
        "input_data": "neighbourhoods_mview",
        "preparation_hook": "neighbourhoods",
        "fitter": "bfs",
        "fitter.ngetter": "lookup",
        "fitter.na": "vuset",
        "fitter.checker": "switch",
        "fitter.queue": "fifo",
    },
    "sorted_neighbourhoods": {
        "input_data": "neighbourhoods_mview",
        "preparation_hook": "neighbourhoods",
        "fitter": "bfs",
        "fitter.ngetter": ("lookup", (), {"is_sorted": True}),
        "fitter.na": "vector",
        "fitter.checker": "screen",
        "fitter.queue": "fifo",
    },
}


def get_registered_recipe(key):
    return dict(REGISTERED_RECIPES[key.lower()])


COMPONENT_ALT_KW_MAP = {
    "prep": "preparation_hook",
    "input": "input_data",
    "data": "input_data",
    "n": "neighbours",
    "na": "neighbours",
    "nb": "neighbour_neighbours",
    "getter": "neighbours_getter",
    "ogetter": "neighbours_getter_other",
    "ngetter": "neighbours_getter",
    "ongetter": "neighbours_getter_other",
    "dgetter": "distance_getter",
    "checker": "similarity_checker",
    "q": "queue",
    "pq": "priority_queue",
    "prioq": "priority_queue",
    "pqt": "priority_queue_tree",
    "prioq_tree": "priority_queue_tree",
    "hfitter": "hierarchical_fitter",
}

COMPONENT_KW_TYPE_ALIAS_MAP = {
    "neighbour_neighbours": "neighbours",
    "neighbour_getter_other": "neighbours_getter",
    # the reference's map spells this key "neighbour_getter_other" (recipes.py:388-392),
    # so a string-named getter for the predictor's other-getter slot is not found there
    "neighbours_getter_other": "neighbours_getter",
    "priority_queue_tree": "priority_queue",
}

COMPONENT_NAME_TYPE_MAP = {
    "preparation_hook": {
        "points_from_parts": prepare_points_from_parts,
        "neighbourhoods": prepare_neighbourhoods,
        "pass": prepare_pass,
    },
    "input_data": {
        "components_mview": _types.InputDataComponents,
        "components": _types.InputDataComponents,
        "neighbourhoods_mview": _types.InputDataNeighbourhoods,
        "neighbourhoods": _types.InputDataNeighbourhoods,
    },
    "neighbours_getter": {
        "brute_force": _types.NeighboursGetterGPU,
        "gpu": _types.NeighboursGetterGPU,
        "lookup": _types.NeighboursGetterLookup,
    },
    "distance_getter": {
        "metric": _types.DistanceGetterMetric,
    },
    "neighbours": {
        "vector": _types.NeighboursCSR,
        "uset": _types.NeighboursCSR,
        "vuset": _types.NeighboursCSR,
        "csr": _types.NeighboursCSR,
    },
    "metric": {
        "euclidean_r": _types.MetricEuclideanReduced,
        "euclidean_reduced": _types.MetricEuclideanReduced,
        "euclidean": _types.MetricEuclidean,
        "precomputed": _types.MetricPrecomputed,
        "euclidean_periodic_r": _types.MetricEuclideanPeriodicReduced,
        "euclidean_periodic_reduced": _types.MetricEuclideanPeriodicReduced,
    },
    "similarity_checker": {
        "contains": _types.SimilarityCheckerGPU,
        "switch": _types.SimilarityCheckerGPU,
        "screen": _types.SimilarityCheckerGPU,
        "gpu": _types.SimilarityCheckerGPU,
        "merge": (lambda: _types.SimilarityCheckerGPU("merge")),
        "bitset": (lambda: _types.SimilarityCheckerGPU("bitset")),
    },
    "queue": {
        "fifo": _types.QueueUnused,
    },
    "fitter": {
        "bfs": _fit.FitterGPU,
        "gpu": _fit.FitterGPU,
    },
    "hierarchical_fitter": {
        "repeat": _fit.HierarchicalFitterRepeat,
        "mst": _fit.HierarchicalFitterMSTPrim,
    },
    "predictor": {
        "firstmatch": _fit.PredictorGPU,
        "gpu": _fit.PredictorGPU,
    },
    "priority_queue": {
        "maxheap": _types.PriorityQueueMaxHeap,
    },
}

# Reference components without a GPU implementation in this engine, with the reason.
UNSUPPORTED_COMPONENTS = {
    "metric": {
        "dummy": "MetricExtDummy is not implemented",
    },
    "distance_getter": {"lookup": "DistanceGetterExtLookup (precomputed distances) is not implemented"},
    "fitter": {"bfs_debug": "FitterExtBFSDebug (step-by-step debug generator) is not implemented"},
}
