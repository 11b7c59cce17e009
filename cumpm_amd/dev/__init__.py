from .electric_field import (DirectElectricField, FieldCalcMode, MonodisperseEwaldElectricField,
                             PolydisperseEwaldElectricField)

__all__ = ["DirectElectricField", "MonodisperseEwaldElectricField", "PolydisperseEwaldElectricField",
           "FieldCalcMode"]
