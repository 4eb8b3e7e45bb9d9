"""Writes nt2_unit_values.json: expected values that NT2's own unit tests hold.

The numbers are data copied from the reference's test files (no code):
  normcdf on -5:1:5   modules/core/statistics/unit/scalar/normcdf.cpp:53-56
  erfc specials       modules/core/euler/unit/scalar/erfc.cpp:45-52
  normpdf             modules/core/statistics/unit/scalar/normpdf.cpp:30-35 (specials),
                      :44-48 (normpdf(a, 1) for a = -5:1:5, i.e. phi(a - 1)), :49-50
                      (normpdf(a, a) = phi(0))
  fastnormcdf         modules/core/statistics/unit/scalar/fastnormcdf.cpp:26-31
Run: python tests/golden/make_nt2_unit_values.py
"""
import json
import os

data = {
    "normcdf": {
        "source": "modules/core/statistics/unit/scalar/normcdf.cpp:53-56 (tolerance 5 NT2-ULP)",
        "x": [-5, -4, -3, -2, -1, 0, 1, 2, 3, 4, 5],
        "y": [2.8665157187948740877e-07, 3.1671241833184224292e-05, 0.0013498980316300926256,
              0.022750131948179170421, 0.15865525393145676891, 0.5, 0.84134474606854325884,
              0.97724986805182079141, 0.99865010196836989653, 0.999968328758166769, 0.99999971334842807646],
    },
    "erfc": {
        "source": "modules/core/euler/unit/scalar/erfc.cpp:45-52 (tolerance 200 NT2-ULP)",
        "x": [0.5, 1.0, 2.0, 6.0, -6.0, 0.0],
        "y": [0.479500122186953, 0.157299207050285, 0.004677734981047, 2.151973671249892e-17, 2.0, 1.0],
    },
    "normpdf": {
        "source": "modules/core/statistics/unit/scalar/normpdf.cpp:30-35 (0.5 NT2-ULP; 0 at zero) "
                  "and :44-50 (1 NT2-ULP); x is the standardised argument a - mu",
        "specials_x": [1.0, -1.0, 0.0],
        "specials_y": [0.24197072451914334979783019293556, 0.24197072451914334979783019293556,
                       0.398942280401432677939946059934],
        "specials_tol": [0.5, 0.5, 0.0],
        "x": [-6, -5, -4, -3, -2, -1, 0, 1, 2, 3, 4],
        "y": [6.0758828498232860873e-09, 1.4867195147342976779e-06, 0.00013383022576488536764,
              0.0044318484119380075273, 0.053990966513188062836, 0.24197072451914336533,
              0.39894228040143270286, 0.24197072451914336533, 0.053990966513188062836,
              0.0044318484119380075273, 0.00013383022576488536764],
        "tol": 1.0,
    },
    "fastnormcdf": {
        "source": "modules/core/statistics/unit/scalar/fastnormcdf.cpp:26-31 (float)",
        "x": [1.0, -1.0, 0.0],
        "y": [0.841344746068542948585232545632, 0.158655253931457051414767454368, 0.5],
        "tol": [1.5, 5.5, 0.5],
    },
}
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "nt2_unit_values.json")
with open(out, "w") as f:
    json.dump(data, f, indent=1)
print("wrote", out)
