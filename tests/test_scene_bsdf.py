"""nlpf_bsdfs_from_scene on the reference's own scene file (tests/golden: the material graphs of
veach_ajar): each material's base surface below its sty_* nodes, and the principled parameter
sets the world filter evaluates exactly (no precomputed albedo table). Needs no GPU."""
import os

import numpy as np
import pytest

import nlpf as N

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "veach_ajar_sty_materials.json")


def _one(name):
    return N.bsdfs_from_scene(GOLDEN, [name])[0]


def test_supported_materials():
    t = N.bsdfs_from_scene(GOLDEN, ["Gold", "Hinge", "Dragon", "MiniLightLeft", "Material_007", "Glass", "Mirror"])
    gold, hinge, dragon, light, m007, glass, mirror = t
    assert gold["kind"] == N.BSDF_PRINCIPLED and gold["metallic"] == 1.0 and gold["roughness"] == 0.0
    assert np.array_equal(gold["specular_tint"], [1, 1, 1]) and gold["specular_ior_level"] == np.float32(0.5)
    assert hinge["kind"] == N.BSDF_PRINCIPLED and hinge["roughness"] == np.float32(0.115175396)
    assert dragon["kind"] == N.BSDF_DIFFUSE
    assert light["kind"] == N.BSDF_EMISSION
    # metallic 0 with no specular layer: the diffuse part of the principled mixture
    assert m007["kind"] == N.BSDF_PRINCIPLED and m007["metallic"] == 0.0 and m007["specular_ior_level"] == 0.0
    # metallic 1 never evaluates the transmission side
    assert glass["kind"] == N.BSDF_PRINCIPLED and glass["transmission"] == 1.0
    assert mirror["roughness"] == np.float32(0.003937006)
    assert not t["reserved"].any()


@pytest.mark.parametrize("name,why", [
    ("BadBook", "specular_ior_level 0"),       # metallic 0 with the default specular layer
    ("Floor", "specular_ior_level 0"),         # metallic 0.24, specular 0.5
    ("Flashlight_Glass", "coat"),              # coat_weight 0.44
    ("Pot3_001", "transmission 0"),            # metallic 0, transmission 1
])
def test_unsupported_principled_sets_name_the_material(name, why):
    with pytest.raises(N.NlpfError) as e:
        _one(name)
    assert e.value.code == 5  # NLPF_E_UNSUPPORTED
    assert name in str(e.value) and why in str(e.value)


def test_errors():
    with pytest.raises(N.NlpfError) as e:
        _one("NoSuchMaterial")
    assert e.value.code == 2 and "NoSuchMaterial" in str(e.value)
    scene = ('{"materials": {"m": {"shader": {"kind": "surface", "output": {"id": "o"}, "nodes": {'
             '"o": {"type": "output", "node": {"id": "g"}}, "g": {"type": "glass"}}}}}}')
    with pytest.raises(N.NlpfError) as e:
        N.bsdfs_from_scene(scene, ["m"])
    assert e.value.code == 5 and "glass" in str(e.value)
    # a principled input driven by another node (not a constant) is rejected
    scene = ('{"materials": {"m": {"shader": {"kind": "surface", "output": {"id": "o"}, "nodes": {'
             '"o": {"type": "output", "node": {"id": "s"}}, '
             '"s": {"type": "sty_toon", "surface": {"id": "p"}}, '
             '"p": {"type": "principled", "metallic": {"id": "img"}}, "img": {"type": "image"}}}}}}')
    with pytest.raises(N.NlpfError) as e:
        N.bsdfs_from_scene(scene, ["m"])
    assert "metallic" in str(e.value)
