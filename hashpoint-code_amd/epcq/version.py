__version__ = "0.3.0+epcq.gfx950"
