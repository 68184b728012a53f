"""The renderer plugin as the reference's machinery builds it (CPU only).

The reference registers renderers by ``__name__`` in a registry that refuses a
second object of the same name (pointrix/utils/registry.py:18-41), and
``parse_renderer`` (pointrix/model/renderer/__init__.py:4-19, fed from
model/base_model.py:43) pops ``name`` from an OmegaConf DictConfig and calls
``cls(cfg, white_bg=..., device=...)``, whose BaseObject.__init__ turns the
mapping into the class's Config (utils/base.py:32-39, parse_structured at
utils/config.py:102-107).  omegaconf is not installed here, so a small
MutableMapping that is not a dict stands in for DictConfig, and a registry
written to the same contract stands in for RENDERER_REGISTRY.

Nothing here launches GPU work: constructing the renderer and the
SH-degree schedule are host-side.
"""
from collections.abc import MutableMapping
from types import SimpleNamespace

import pytest

from pointrix_amd import renderer as R


class StandInRegistry:
    """The reference Registry's contract: register under obj.__name__, assert
    the name is new, KeyError from get for an unknown name."""

    def __init__(self, name):
        self._name, self._obj_map = name, {}

    def _do_register(self, name, obj):
        assert name not in self._obj_map, \
            f"An object named '{name}' was already registered in '{self._name}' registry!"
        self._obj_map[name] = obj

    def register(self, obj=None):
        if obj is None:
            def deco(o):
                self._do_register(o.__name__, o)
                return o
            return deco
        self._do_register(obj.__name__, obj)

    def get(self, name):
        ret = self._obj_map.get(name)
        if ret is None:
            raise KeyError(f"No object named '{name}' found in '{self._name}' registry!")
        return ret


class ConfigMapping(MutableMapping):
    """A mapping that is not a dict (the role of OmegaConf's DictConfig)."""

    def __init__(self, **kw):
        self._d = dict(kw)

    def __getitem__(self, k):
        return self._d[k]

    def __setitem__(self, k, v):
        self._d[k] = v

    def __delitem__(self, k):
        del self._d[k]

    def __iter__(self):
        return iter(self._d)

    def __len__(self):
        return len(self._d)


def reference_parse_renderer(registry, cfg, **kwargs):
    """parse_renderer's two statements for a name outside its special cases."""
    name = cfg.pop("name")
    return registry.get(name)(cfg, **kwargs)


def fresh_registry():
    reg = StandInRegistry("RENDERER")

    @reg.register()
    class MsplatRender:                     # the reference's own class, by name
        def __init__(self, cfg, **kw):
            self.cfg = cfg

    return reg


def test_integration_snippet_registers_and_builds():
    """The registration INTEGRATION.md documents, then the reference's
    parse_renderer path with a non-dict mapping config."""
    reg = fresh_registry()

    @reg.register()
    class AmdMsplatRender(R.MsplatRender):
        pass

    cfg = ConfigMapping(name="AmdMsplatRender", update_sh_iter=500, max_sh_degree=2,
                        render_depth=True)
    r = reference_parse_renderer(reg, cfg, white_bg=True, device="cuda")
    assert isinstance(r, R.MsplatRender) and type(r).__name__ == "AmdMsplatRender"
    assert isinstance(r.cfg, R.MsplatRender.Config)
    assert (r.cfg.update_sh_iter, r.cfg.max_sh_degree, r.cfg.render_depth) == (500, 2, True)
    # the build's own fields keep their defaults
    assert r.cfg.fused is True and r.cfg.tight_tiles is False and r.cfg.render_normal is False
    assert r.bg_color == 1.0 and r.sh_degree == 0 and r.device == "cuda"
    assert "name" not in cfg                 # parse_renderer popped it


def test_same_name_collides_in_the_reference_registry():
    """Registering the class under its own name collides with the reference's
    MsplatRender: the reason the integration registers a subclass."""
    reg = fresh_registry()
    with pytest.raises(AssertionError):
        reg.register(R.MsplatRender)


def test_own_registry_has_the_reference_contract():
    reg = R.Registry("RENDERER")

    @reg.register()
    class A:
        pass

    with pytest.raises(AssertionError):
        reg.register(A)
    with pytest.raises(KeyError):
        reg.get("B")
    assert reg.get("A") is A and "A" in reg and dict(reg) == {"A": A}
    assert "MsplatRender" in R.RENDERER_REGISTRY and "MsplatNormalRender" in R.RENDERER_REGISTRY
    r = R.parse_renderer(ConfigMapping(name="MsplatNormalRender", render_depth=True))
    assert isinstance(r, R.MsplatNormalRender) and r.cfg.render_depth is True


@pytest.mark.parametrize("cfg", [None, {}, ConfigMapping(),
                                 SimpleNamespace(max_sh_degree=1, render_depth=True),
                                 R.MsplatRender.Config(max_sh_degree=1, render_depth=True)])
def test_config_forms(cfg):
    r = R.MsplatRender(cfg, white_bg=False, device="cuda")
    assert isinstance(r.cfg, R.MsplatRender.Config) and r.bg_color == 0.0
    want = (1, True) if isinstance(cfg, (SimpleNamespace, R.MsplatRender.Config)) else (3, False)
    assert (r.cfg.max_sh_degree, r.cfg.render_depth) == want


def test_unknown_key_raises_like_the_reference():
    # parse_structured builds Config(**cfg): an unknown field is a TypeError
    with pytest.raises(TypeError):
        R.MsplatRender(ConfigMapping(render_dpeth=True))


def test_sh_degree_schedule_and_state_dict():
    """update_sh_degree / state_dict / load_state_dict (msplat.py:215-247):
    +1 every update_sh_iter steps up to max_sh_degree; the checkpoint holds
    {"sh_degree": d} and restores it."""
    r = R.MsplatRender({"update_sh_iter": 3, "max_sh_degree": 2})
    seen = []
    for step in range(1, 13):
        r.update_sh_degree(step)
        seen.append(r.sh_degree)
    assert seen == [0, 0, 1, 1, 1, 2, 2, 2, 2, 2, 2, 2]
    sd = r.state_dict()
    assert sd == {"sh_degree": 2}
    r2 = R.MsplatRender({"update_sh_iter": 3, "max_sh_degree": 2})
    assert r2.sh_degree == 0
    r2.load_state_dict(sd)
    assert r2.sh_degree == 2 and r2.state_dict() == sd
