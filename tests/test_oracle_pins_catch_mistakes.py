"""CPU: every pin of the oracle rejects a plausible mistake planted in the
oracle.

Each test plants one mistake in the oracle on the CPU (a transposed operand, a
swapped index pair, a dropped conjugate, a dropped or sign-flipped term, a
relative cutoff in place of the absolute one, an off-by-one in the random
stream, a conjugated flux phase, spin-1/2 operators in place of Pauli ones) by
monkeypatching, then runs a pin from test_oracle_dmrg.py or
test_oracle_models.py unchanged and requires it to fail.  The unmutated pins
pass in those modules.
"""
import numpy as np
import pytest

import test_oracle_dmrg as dmrg_pins
import test_oracle_models as model_pins
from oracle import dmrg as od
from oracle import models, prng


def _caught(pin, *args):
    """True when the pin fails (an assertion, or an error a mistake causes)."""
    try:
        pin(*args)
    except (AssertionError, ValueError, np.linalg.LinAlgError):
        return True
    return False


class _NoConj(np.ndarray):
    """An array whose conj() is itself: plants a dropped complex conjugate."""

    def conj(self):
        return self


def test_transposed_mpo_operand_in_heff(monkeypatch):
    orig = od.apply_heff2
    monkeypatch.setattr(od, "apply_heff2",
                        lambda Lg, Rg, W1, W2, th: orig(Lg, Rg, W1.transpose(0, 2, 1, 3), W2, th))
    assert _caught(dmrg_pins.test_apply_heff2_matches_reference_einsum, True)


def test_ket_bra_swapped_environment_in_heff(monkeypatch):
    orig = od.apply_heff2
    monkeypatch.setattr(od, "apply_heff2",
                        lambda Lg, Rg, W1, W2, th: orig(Lg.transpose(2, 1, 0), Rg, W1, W2, th))
    assert _caught(dmrg_pins.test_apply_heff2_matches_reference_einsum, False)


@pytest.mark.parametrize("side", ["left", "right"])
def test_dropped_conjugate_in_environment_update(monkeypatch, side):
    name = "update_%s_env" % side
    orig = getattr(od, name)
    monkeypatch.setattr(od, name, lambda E, A, W: orig(E, np.asarray(A).view(_NoConj), W))
    assert _caught(dmrg_pins.test_update_envs_match_reference_einsums, True)
    # control: on real tensors conj() is the identity, and the pin passes
    assert not _caught(dmrg_pins.test_update_envs_match_reference_einsums, False)


def test_swapped_mpo_bond_indices_in_environment_update(monkeypatch):
    orig = od.update_right_env
    monkeypatch.setattr(od, "update_right_env",
                        lambda R, A, W: orig(R, A, W.transpose(3, 1, 2, 0)))
    assert _caught(dmrg_pins.test_update_envs_match_reference_einsums, False)


def _off_diagonal_scaled(Ws, f, every=1):
    """Scale the entries with s != sp (the S+ and S- factors of the hopping
    terms) by f on every `every`-th site."""
    out = []
    for i, W in enumerate(Ws):
        W = W.copy()
        if i % every:
            out.append(W)
            continue
        d = W.shape[1]
        for s in range(d):
            for sp in range(d):
                if s != sp:
                    W[:, s, sp, :] *= f
        out.append(W)
    return out


def test_sign_flipped_hopping_in_mpo(monkeypatch):
    """A flipped hopping sign (one factor of each term negated: every other
    site) is unitarily equivalent on a chain, by a rotation of every other
    site, so energy checks alone pass it; the element-wise MPO = Kronecker
    pin does not."""
    orig = models.heisenberg_mpo
    monkeypatch.setattr(models, "heisenberg_mpo",
                        lambda L, J=1.0: _off_diagonal_scaled(orig(L, J), -1.0, every=2))
    E_mut = models.ground_energy(models.mpo_to_dense(models.heisenberg_mpo(4)))
    assert abs(E_mut - (-1.616025403784)) < 1e-10   # the spectrum does not see it
    assert _caught(model_pins.test_heisenberg_mpo_equals_kronecker, 4)


def test_dropped_hopping_in_mpo(monkeypatch):
    orig = models.heisenberg_mpo
    monkeypatch.setattr(models, "heisenberg_mpo", lambda L, J=1.0: _off_diagonal_scaled(orig(L, J), 0.0))
    assert _caught(model_pins.test_heisenberg_mpo_equals_kronecker, 4)
    assert _caught(dmrg_pins.test_dmrg2_known_answers, 4, -1.616025403784)


def test_relative_instead_of_absolute_cutoff(monkeypatch):
    orig = od.truncation_rank
    monkeypatch.setattr(od, "truncation_rank",
                        lambda S, k, cutoff=od.SVD_ABS_CUTOFF, rel_cutoff=0.0:
                        orig(S, k, 0.0, max(cutoff, rel_cutoff)))
    assert _caught(dmrg_pins.test_svd_split2_truncation_rule)


def test_off_by_one_random_stream(monkeypatch):
    orig = prng.splitmix64
    monkeypatch.setattr(prng, "splitmix64", lambda seed, count, offset=0: orig(seed, count, offset + 1))
    assert _caught(model_pins.test_splitmix64_published_vector)


def test_conjugated_flux_phase_in_josephson_mpo(monkeypatch):
    """The wrong sign of the flux phase gives conj(H): the same spectrum, so
    the known answers pass it; the element-wise MPO = Kronecker pin does not."""
    orig = models.josephson_mpo

    def mutant(L, n_max=2, E_J=2.0, E_C=1.0, phi_ext=np.pi / 3, n_g=0.0):
        return orig(L, n_max=n_max, E_J=E_J, E_C=E_C, phi_ext=-phi_ext, n_g=n_g)

    monkeypatch.setattr(models, "josephson_mpo", mutant)
    kw = dict(n_max=1, E_J=1.0, E_C=0.5, phi_ext=np.pi / 4)
    E_mut = models.ground_energy(models.mpo_to_dense(models.josephson_mpo(4, **kw)))
    assert abs(E_mut - (-1.053346829927396)) < 1e-10   # the spectrum does not see it
    assert _caught(model_pins.test_josephson_mpo_equals_kronecker, 3)
    # and the reference's own Josephson MPO fixes which sign is right
    assert _caught(model_pins.test_fixture_josephson_mpo_matches_builder)


def test_spin_half_operators_in_tfim(monkeypatch):
    """S = sigma/2 in place of the Pauli matrices scales the bond term by 1/4
    and the field by 1/2: the free-fermion closed form rejects it."""
    orig = models.dense_tfim
    monkeypatch.setattr(models, "dense_tfim", lambda L, J=1.0, hx=1.0: orig(L, J / 4, hx / 2))
    assert _caught(model_pins.test_tfim_ed_free_fermion_closed_form, 5, 1.0, 0.3)


def test_flipped_field_sign_in_tfim_mpo(monkeypatch):
    """-hx for hx is unitarily equivalent (rotate every site about z), so only
    the element-wise MPO = Kronecker pin sees it."""
    orig = models.tfim_mpo
    monkeypatch.setattr(models, "tfim_mpo", lambda L, J=1.0, hx=1.0: orig(L, J, -hx))
    E_mut = models.ground_energy(models.mpo_to_dense(models.tfim_mpo(5, J=0.8, hx=0.6)))
    E_ok = models.ground_energy(models.dense_tfim(5, J=0.8, hx=0.6))
    assert abs(E_mut - E_ok) < 1e-12          # the spectrum does not see it
    assert _caught(model_pins.test_tfim_mpo_equals_kronecker, 5)


def test_transposed_mpo_operand_in_single_site_heff(monkeypatch):
    """For a Hermitian MPO a transposed operand gives conj(H_eff), still
    Hermitian; the identity-site relation to the einsum-pinned two-site H_eff
    (random, non-Hermitian tensors) rejects it."""
    orig = od.apply_heff1
    monkeypatch.setattr(od, "apply_heff1",
                        lambda Lg, Rg, W, th: orig(Lg, Rg, W.transpose(0, 2, 1, 3), th))
    assert not _caught(dmrg_pins.test_heff_is_hermitian_for_a_hermitian_mpo, False)
    assert _caught(dmrg_pins.test_apply_heff1_is_heff2_with_an_identity_site, True)
