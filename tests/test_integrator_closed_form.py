"""The oracle's integrator end to end against answers the mathematics fixes:
direct lighting through the full render for every MIS heuristic and light
picking strategy (with the reference's emission weight and shadow-origin
rules, see integrator_cases.py), the textbook estimate failing where those
rules matter (so the check is sensitive to them), and the balance / power
heuristics converging to the same image on a multi-light Cornell box.  The
GPU runs the same cases in test_gpu_integrator.py."""
import pytest

import integrator_cases as IC

ESTIMATORS = [(m, p) for m in ("balance", "power") for p in ("all", "power", "uniform")]


def _oracle_fns(oracle):
    def render(desc, st):
        return oracle.OracleScene(desc).render(st)["beauty"]

    def primary(desc, st):
        p, t, u, v, _, _ = oracle.OracleScene(desc).trace_primary(st, 0, st.samples_per_pixel)
        return p, t, u, v
    return render, primary


@pytest.mark.parametrize("emitters", [("sphere",), ("quad",), ("area",)])
@pytest.mark.parametrize("mis", ["balance", "power"])
def test_direct_lighting_single_emitter(oracle, emitters, mis):
    render, primary = _oracle_fns(oracle)
    IC.check_direct_lighting(render, primary, emitters, mis, "all")


@pytest.mark.parametrize("mis,pick", ESTIMATORS)
def test_direct_lighting_three_emitters(oracle, mis, pick):
    render, primary = _oracle_fns(oracle)
    IC.check_direct_lighting(render, primary, ("sphere", "quad", "area"), mis, pick)


@pytest.mark.parametrize("emitters,pick", [(("sphere",), "all"), (("quad", "area"), "uniform")])
def test_textbook_estimate_is_rejected(oracle, emitters, pick):
    """Without the self-shadowing of grazing samples (sphere), or with the
    textbook emission weight under light picking (planar emitters, which
    have no self-shadowing here), rho / pi * E is off by far more than the
    tolerance: the pin sees both rules."""
    render, primary = _oracle_fns(oracle)
    mean_d, se, total, _, _ = IC.direct_lighting_case(render, primary, emitters, "balance", pick, spp=64, runs=8,
                                                      model=False)
    print(f"textbook {emitters} {pick}: relative bias {mean_d / total:+.3e} (3 se {3 * se / total:.1e})")
    assert abs(mean_d) > 5 * (3 * se + 2e-4 * total)


def test_planar_emitters_avoid_the_self_shadow_rule():
    assert IC.min_planar_cos() > 0.7 > (5 ** 0.5 - 1) / 2


def test_estimator_invariance_cornell(oracle):
    render, _ = _oracle_fns(oracle)
    IC.check_estimator_invariance(IC.estimator_means(render))


@pytest.mark.parametrize("mis,pick", [("balance", "all"), ("power", "all"), ("balance", "uniform")])
def test_sky_furnace(oracle, mis, pick):
    render, primary = _oracle_fns(oracle)
    IC.check_sky_furnace(render, primary, mis, pick)


def test_sky_environment_light_record(oracle):
    """SceneLoader adds the EnvironmentLight last, with the override's shadow
    samples or one; a black sky adds none."""
    from raytracer import abi
    d = IC.sky_scene(8, 8, extra_light=True)
    assert list(d.lights["kind"]) == [abi.RT_LIGHT_SPHERE, abi.RT_LIGHT_ENVIRONMENT]
    assert d.lights[1]["shadow_samples"] == 1
    assert len(IC.sky_scene(8, 8, sky=(0.0, 0.0, 0.0)).lights) == 0
