"""Typed inputs of ``simus()`` / ``pfield()``: transducer, baffle and medium.

Field names, defaults and validation rules follow the reference models
(``transducer_params.py:10-160``, ``medium_params.py:6-22``) so that keyword
construction and the raised errors are the same. ``BaffleType`` is a ``str``
Enum because ``enum.StrEnum`` does not exist on Python 3.10.
"""

from __future__ import annotations

from enum import Enum
from math import inf

from pydantic import BaseModel, ConfigDict, Field, NonNegativeFloat, model_validator


class BaffleType(str, Enum):
    """Acoustic boundary condition around each element.

    SOFT: pressure release, obliquity factor cos(theta).
    RIGID: pressure doubling, obliquity factor 1.
    A non-negative float baffle value instead selects the impedance-ratio
    model cos(theta) / (cos(theta) + value).
    """

    SOFT = "soft"
    RIGID = "rigid"

    def __str__(self) -> str:
        return self.value


class TransducerParams(BaseModel):
    """Linear or convex array description (SI units)."""

    model_config = ConfigDict(frozen=True)

    freq_center: float = Field(..., gt=0)
    pitch: float = Field(..., gt=0)
    n_elements: int = Field(..., gt=0)
    width: float | None = Field(None, gt=0)
    kerf: float | None = Field(None, ge=0)
    height: float = Field(default=inf, gt=0)
    elev_focus: float = Field(default=inf, gt=0)
    radius: float = Field(default=inf, gt=0)
    bandwidth: float = Field(default=0.75, gt=0, le=2.0)
    baffle: BaffleType | NonNegativeFloat = Field(default=BaffleType.SOFT)

    @property
    def element_width(self) -> float:
        """Active element width: ``width`` if given, else ``pitch - kerf``."""
        if self.width is not None:
            return self.width
        if self.kerf is not None:
            return self.pitch - self.kerf
        raise ValueError("Either width or kerf must be provided")

    @property
    def kerf_width(self) -> float:
        """Gap between elements: ``kerf`` if given, else ``pitch - width``."""
        if self.kerf is not None:
            return self.kerf
        if self.width is not None:
            return self.pitch - self.width
        raise ValueError("Either width or kerf must be provided")

    @model_validator(mode="after")
    def _check_width_kerf(self) -> "TransducerParams":
        if self.width is None and self.kerf is None:
            raise ValueError("Either width or kerf must be provided")
        if self.width is not None and self.kerf is not None:
            raise ValueError("Cannot specify both width and kerf. Provide only one.")
        if self.width is not None and self.width > self.pitch:
            raise ValueError(f"Element width ({self.width}) cannot exceed pitch ({self.pitch})")
        if self.kerf is not None and self.kerf >= self.pitch:
            raise ValueError(f"Kerf ({self.kerf}) must be less than pitch ({self.pitch})")
        return self


class MediumParams(BaseModel):
    """Propagation medium: speed of sound (m/s), attenuation (dB/cm/MHz)."""

    model_config = ConfigDict(frozen=True)

    speed_of_sound: float = Field(default=1540.0, gt=0)
    attenuation: float = Field(default=0.0, ge=0)


def baffle_code(baffle: BaffleType | float) -> tuple[int, float]:
    """Encode a baffle for the C ABI: (0, 0) soft, (1, 0) rigid, (2, b) impedance."""
    if isinstance(baffle, BaffleType):
        return (0, 0.0) if baffle == BaffleType.SOFT else (1, 0.0)
    if isinstance(baffle, str):
        return baffle_code(BaffleType(baffle.lower()))
    return 2, float(baffle)
