# cfg of the Q10 plan, keyed by opId in construction order:
# 3 = the orders date filter, 8 = the lineitem return-flag filter,
# 6 = the nation join, 9 = the lineitem join

{ 3: { "selectivity": 0.05 },
  8: { "selectivity": 0.34 },
  6: { "multimatch": False },
  9: { "selectivity": 0.034, "multimatch": False } }
