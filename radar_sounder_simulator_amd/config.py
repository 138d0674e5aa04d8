"""Run configuration (pydantic), the fields and validators of the reference's
``config.py:8-377`` that the coherent hot path consumes.

Same model names, field names, defaults and ``ValueError`` cases: waveform
chirp parameters, antenna pattern parameters, media/interface count, roughness
and grazing-fix coherent-only rules, exponential ACF requiring the grazing fix,
offset-interface references.  Extensions: ``WaveformConfig.window`` also
accepts ``"tukey"`` (with ``tukey_alpha``); ``SimConfig.precision`` picks the
device tier ("mixed" or "fp64"); ``SimConfig.joint_form`` picks the joint
solver: "reference" (default) the reference's 2N-unknown damped Newton,
"flat_stack" its two-unknown form for stacks whose interfaces below the
surface are horizontal planes (opt-in; it converges where the reference's
budget can stop short).
"""

from typing import Annotated, Literal, Optional, Union

from pydantic import BaseModel, Field, model_validator


class WaveformConfig(BaseModel):
    kind: Literal["delta", "chirp"] = "delta"
    bandwidth: Optional[float] = None
    pulse_length: Optional[float] = None
    window: Literal["none", "hann", "hamming", "tukey"] = "hann"
    tukey_alpha: float = 0.2
    construction: Literal["analytic", "chirp"] = "analytic"
    interp_bins: bool = False
    incoherent_envelope: bool = False

    @model_validator(mode="after")
    def _chirp_params(self):
        if self.kind == "chirp":
            if not self.bandwidth or self.bandwidth <= 0:
                raise ValueError("chirp waveform requires bandwidth > 0")
            if not self.pulse_length or self.pulse_length <= 0:
                raise ValueError("chirp waveform requires pulse_length > 0")
        if not (0.0 <= self.tukey_alpha <= 1.0):
            raise ValueError("tukey_alpha must be in [0, 1]")
        return self


class AntennaConfig(BaseModel):
    kind: Literal["isotropic", "dipole", "array", "array_tapered", "finite_dipole",
                  "tabulated"] = "isotropic"
    axis: Literal["along_track", "cross_track"] = "along_track"
    n_elements: int = 5
    spacing_lam: float = 0.5
    element_directivity_db: Optional[float] = Field(default=None, ge=3.0,
                                                    allow_inf_nan=False)
    tx_weights: Optional[list[float]] = None
    rx_weights: Optional[list[float]] = None
    length_lam: float = 0.5
    theta_deg: Optional[list[float]] = None
    gain: Optional[list[float]] = None
    roll_source: Literal["none", "nav"] = "none"

    @model_validator(mode="after")
    def _pattern_params(self):
        if self.element_directivity_db is not None and self.kind not in (
                "array", "array_tapered"):
            raise ValueError("element_directivity_db is only valid for array antennas")
        if self.kind == "array":
            if self.n_elements < 2:
                raise ValueError("array antenna requires n_elements >= 2")
            if self.spacing_lam <= 0:
                raise ValueError("array antenna requires spacing_lam > 0")
        if self.kind == "array_tapered":
            if not self.tx_weights or not self.rx_weights:
                raise ValueError("array_tapered antenna requires tx_weights and rx_weights")
            if len(self.tx_weights) != len(self.rx_weights) or len(self.tx_weights) < 2:
                raise ValueError("array_tapered needs >= 2 matching tx_weights/rx_weights")
            for w in (self.tx_weights, self.rx_weights):
                if any(v < 0 for v in w) or sum(w) <= 0:
                    raise ValueError("array_tapered weights must be >= 0 with a positive sum")
            if self.spacing_lam <= 0:
                raise ValueError("array_tapered requires spacing_lam > 0")
        if self.kind == "finite_dipole" and not (0.0 < self.length_lam <= 1.0):
            raise ValueError("finite_dipole requires 0 < length_lam <= 1 (carrier wavelengths)")
        if self.kind == "tabulated":
            if not self.theta_deg or not self.gain:
                raise ValueError("tabulated antenna requires theta_deg and gain samples")
            if len(self.theta_deg) != len(self.gain) or len(self.gain) < 2:
                raise ValueError("tabulated antenna needs >= 2 matching theta_deg/gain")
            t = self.theta_deg
            if any(b <= a for a, b in zip(t, t[1:])):
                raise ValueError("tabulated theta_deg must be ascending")
            if any(g < 0 for g in self.gain):
                raise ValueError("tabulated gain samples must be >= 0 (one-way FIELD gain)")
        return self


class RadarConfig(BaseModel):
    dt: float
    n_samples: int
    t0: float
    c: float = 299792458.0
    f0: Optional[float] = None
    waveform: WaveformConfig = Field(default_factory=WaveformConfig)
    antenna: AntennaConfig = Field(default_factory=AntennaConfig)

    @property
    def wavelength(self) -> float:
        if self.f0 is None:
            raise ValueError("wavelength requires radar.f0 to be set")
        return self.c / self.f0


class Medium(BaseModel):
    name: str
    eps_r: float
    attenuation_db_per_km: float = 0.0


class FacetConfig(BaseModel):
    spacing: Optional[float] = None


class RoughnessConfig(BaseModel):
    sigma_m: float
    corr_length_m: float
    acf: Literal["gaussian", "exponential"] = "gaussian"

    @model_validator(mode="after")
    def _positive(self):
        if self.sigma_m < 0:
            raise ValueError("roughness sigma_m must be >= 0")
        if self.corr_length_m <= 0:
            raise ValueError("roughness corr_length_m must be > 0")
        return self


class GrazingFixConfig(BaseModel):
    s_eff: float = 0.05

    @model_validator(mode="after")
    def _positive(self):
        if self.s_eff <= 0:
            raise ValueError("grazing_fix s_eff must be > 0")
        return self


class _InterfaceBase(BaseModel):
    name: Optional[str] = None
    roughness: Optional[RoughnessConfig] = None


class DemInterface(_InterfaceBase):
    kind: Literal["dem"] = "dem"
    path: Optional[str] = None
    ref: Optional[str] = None

    @model_validator(mode="after")
    def _one_source(self):
        if self.path is not None and self.ref is not None:
            raise ValueError("dem interface: set at most one of path/ref")
        return self


class FlatInterface(_InterfaceBase):
    kind: Literal["flat"] = "flat"
    elevation: float


class OffsetInterface(_InterfaceBase):
    kind: Literal["offset"] = "offset"
    reference: Union[int, str]
    offset: float


InterfaceConfig = Annotated[Union[DemInterface, FlatInterface, OffsetInterface],
                            Field(discriminator="kind")]


def _default_media():
    return [Medium(name="air", eps_r=1.0), Medium(name="ice", eps_r=3.17)]


def _default_interfaces():
    return [DemInterface(name="surface")]


class SimConfig(BaseModel):
    mode: Literal["incoherent", "coherent"]
    split_sides: bool = False
    refraction: Literal["sequential", "joint"] = "joint"
    grazing_fix: Optional[GrazingFixConfig] = None
    roughness_seed: int = 0
    diffuse_exponent: float = 1.0
    precision: Literal["mixed", "fp64"] = "mixed"
    joint_form: Literal["reference", "flat_stack"] = "reference"
    radar: RadarConfig
    facets: FacetConfig = Field(default_factory=FacetConfig)
    media: list[Medium] = Field(default_factory=_default_media)
    interfaces: list[InterfaceConfig] = Field(default_factory=_default_interfaces)

    @model_validator(mode="after")
    def _check_stack(self):
        if len(self.media) != len(self.interfaces) + 1:
            raise ValueError(
                f"need len(media) == len(interfaces) + 1; got {len(self.media)} media, "
                f"{len(self.interfaces)} interfaces")
        if self.mode != "coherent" and any(i.roughness is not None for i in self.interfaces):
            raise ValueError("interface roughness requires coherent mode")
        if self.mode != "coherent" and self.grazing_fix is not None:
            raise ValueError("grazing_fix requires coherent mode")
        if self.grazing_fix is None and any(
                i.roughness is not None and i.roughness.acf == "exponential"
                for i in self.interfaces):
            raise ValueError("roughness acf='exponential' requires grazing_fix")
        names = [i.name for i in self.interfaces]
        for i, iface in enumerate(self.interfaces):
            if isinstance(iface, OffsetInterface):
                ref = iface.reference
                if isinstance(ref, int):
                    if not 0 <= ref < len(self.interfaces) or ref == i:
                        raise ValueError(f"offset interface {i}: bad reference index {ref}")
                elif ref not in names or names.index(ref) == i:
                    raise ValueError(f"offset interface {i}: unknown reference name {ref!r}")
        return self
