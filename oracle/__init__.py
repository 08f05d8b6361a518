'''The float64 CPU restatement of the delta-SPH step: test infrastructure, the checker and the CPU baseline only.'''
