"""System description for the integral -> H(k)/S(k) -> eigenvalue path.

`SystemInput` holds what the reference reads from its input files before any
derived setup runs (src/uolcao/input.f90:116-244): the cell vectors
(`readRealCellVectors`, lattice.f90:131), atom types with their Gaussian
exponents and radial coefficients (`readAtomicTypes`/`readRadialFns`,
atomicTypes.f90:132,277), atom sites (atomicSites.f90:85), potential types
(potTypes.f90:86, even-tempered alphas from `makePotAlphas` :170) and sites
(potSites.f90:40), the k-points in fractional reciprocal coordinates
(kpoints.f90:141-318, 718-934), the basis-function cutoff
(`setCutoffThresh`, lattice.f90:239) and the potential coefficients
(`initPotCoeffs`, potential.f90:433).

Everything derived from it (basis with angular normalisation, the sorted
real-space cell list, Cartesian k-points, screening, work lists) is computed
by the library itself.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List

import numpy as np

BOHR_PER_ANGSTROM = 1.0 / 0.529177210903   # src/constants.f90 bohrRad


@dataclass
class AtomType:
    element: int                 # element id; screening tables are per element
    alphas: np.ndarray           # ascending Gaussian exponents (bohr^-2)
    num_orb_alphas: tuple        # alphas used by l = 0,1,2,3 (NUM_ALPHA_S_P_D_F)
    vale_l: List[int]            # l of each valence radial function
    vale_coef: List[np.ndarray]  # radial coefficients, one per used alpha
    core_l: List[int] = field(default_factory=list)
    core_coef: List[np.ndarray] = field(default_factory=list)

    @property
    def num_vale_states(self) -> int:
        return sum(2 * l + 1 for l in self.vale_l)

    @property
    def num_core_states(self) -> int:
        return sum(2 * l + 1 for l in self.core_l)


@dataclass
class PotType:
    element: int
    Z: float                     # nuclear charge
    nuc_alpha: float             # NUCLEAR_CHARGE__ALPHA
    alphas: np.ndarray           # ascending s-Gaussian exponents


@dataclass
class SystemInput:
    name: str
    cell: np.ndarray             # (3,3) rows are a_1, a_2, a_3 in bohr
    atom_pos: np.ndarray         # (nAtoms, 3) Cartesian bohr
    atom_type: np.ndarray        # (nAtoms,) 0-based index into types
    types: List[AtomType]
    pot_types: List[PotType]
    pot_pos: np.ndarray          # (nPotSites, 3)
    pot_type: np.ndarray         # (nPotSites,) 0-based, sites grouped by type
    kpoints: np.ndarray          # (nk, 3) fractional reciprocal coordinates
    kweights: np.ndarray         # (nk,)
    pot_coeffs: np.ndarray       # (potDim,) or (spin, potDim)
    basis_cutoff: float = 1.0e-16
    gamma: bool = False          # real Gamma-point build (-DGAMMA)
    rel: int = 0                 # 1: scalar relativistic (mass-velocity term)
    spin: int = 1                # potential-coefficient sets (potCoeffs(:,spin))

    # ---- sizes -----------------------------------------------------------
    @property
    def n_atoms(self) -> int:
        return int(self.atom_pos.shape[0])

    @property
    def vale_dim(self) -> int:
        return int(sum(self.types[t].num_vale_states for t in self.atom_type))

    @property
    def core_dim(self) -> int:
        return int(sum(self.types[t].num_core_states for t in self.atom_type))

    @property
    def pot_dim(self) -> int:
        return int(sum(len(p.alphas) for p in self.pot_types))

    @property
    def nk(self) -> int:
        return 1 if self.gamma else int(self.kpoints.shape[0])

    def validate(self) -> None:
        for t in self.types:
            a = np.asarray(t.alphas)
            if np.any(np.diff(a) <= 0):
                raise ValueError("atomic alphas must be strictly ascending")
            if len(t.vale_l) != len(t.vale_coef) or len(t.core_l) != len(t.core_coef):
                raise ValueError("radial function lists are inconsistent")
        for p in self.pot_types:
            if np.any(np.diff(np.asarray(p.alphas)) <= 0):
                raise ValueError("potential alphas must be strictly ascending")
        if np.any(np.diff(self.pot_type) < 0):
            raise ValueError("potential sites must be grouped by type")
        if self.spin not in (1, 2) or self.rel not in (0, 1):
            raise ValueError("spin must be 1 or 2 and rel 0 or 1")
        if np.asarray(self.pot_coeffs).size != self.pot_dim * self.spin:
            raise ValueError("pot_coeffs must hold spin x potDim values")

    # ---- flat arrays shared by the C interfaces -----------------------------
    def flat(self) -> dict:
        """Contiguous arrays in the layout of the C system descriptors."""
        self.validate()
        nT = len(self.types)
        maxA = max(len(t.alphas) for t in self.types)
        maxR = max(max(len(t.vale_l), len(t.core_l), 1) for t in self.types)
        alphas = np.zeros((nT, maxA))
        nalpha = np.zeros(nT, np.int32)
        norb = np.zeros((nT, 4), np.int32)
        elem = np.zeros(nT, np.int32)
        nvr = np.zeros(nT, np.int32)
        ncr = np.zeros(nT, np.int32)
        vl = np.zeros((nT, maxR), np.int32)
        cl = np.zeros((nT, maxR), np.int32)
        vc = np.zeros((nT, maxR, maxA))
        cc = np.zeros((nT, maxR, maxA))
        for i, t in enumerate(self.types):
            na = len(t.alphas)
            alphas[i, :na] = t.alphas
            nalpha[i] = na
            norb[i, :] = t.num_orb_alphas
            elem[i] = t.element
            nvr[i] = len(t.vale_l)
            ncr[i] = len(t.core_l)
            for r, (l, c) in enumerate(zip(t.vale_l, t.vale_coef)):
                vl[i, r] = l
                vc[i, r, :len(c)] = c
            for r, (l, c) in enumerate(zip(t.core_l, t.core_coef)):
                cl[i, r] = l
                cc[i, r, :len(c)] = c
        nP = len(self.pot_types)
        maxPA = max(len(p.alphas) for p in self.pot_types)
        palpha = np.zeros((nP, maxPA))
        pna = np.zeros(nP, np.int32)
        pZ = np.zeros(nP)
        pnuc = np.zeros(nP)
        pelem = np.zeros(nP, np.int32)
        for i, p in enumerate(self.pot_types):
            palpha[i, :len(p.alphas)] = p.alphas
            pna[i] = len(p.alphas)
            pZ[i] = p.Z
            pnuc[i] = p.nuc_alpha
            pelem[i] = p.element
        c = np.ascontiguousarray
        return dict(
            nAtoms=self.n_atoms, atomPos=c(self.atom_pos, dtype=np.float64),
            atomType=c(self.atom_type, dtype=np.int32),
            nTypes=nT, maxAlphas=maxA, maxRadial=maxR, typeNumAlphas=nalpha,
            typeAlphas=alphas, typeNumOrbAlphas=norb, typeElement=elem,
            typeNumValeRadial=nvr, typeNumCoreRadial=ncr, typeValeL=vl,
            typeCoreL=cl, typeValeCoef=vc, typeCoreCoef=cc,
            nPotTypes=nP, maxPotAlphas=maxPA, potNumAlphas=pna,
            potAlphas=palpha, potZ=pZ, potNucAlpha=pnuc, potElement=pelem,
            nPotSites=int(self.pot_pos.shape[0]),
            potPos=c(self.pot_pos, dtype=np.float64),
            potSiteType=c(self.pot_type, dtype=np.int32),
            cellVectors=c(self.cell, dtype=np.float64).reshape(9),
            nk=int(self.kpoints.shape[0]),
            kpoints=c(self.kpoints, dtype=np.float64),
            basisCutoff=float(self.basis_cutoff), gamma=int(self.gamma),
            rel=int(self.rel), spin=int(self.spin),
        )
