"""Option parameter objects (interface of FiniteDifference/Option.py and the
Heston option classes of FiniteDifference/heston.py:37-271).

Only the attributes the engines read are kept: spot, strike, tenor,
interest_rate.value, variance.{value, mean, reversion, volatility},
correlation and, for barrier options, ``top`` / ``bottom`` as
(knock_in, level) tuples.  Monte Carlo and analytic pricers are not part of
the hot path.
"""


class MeanRevertingProcess(object):
    """Option.py:133-161."""

    attrs = ['mean', 'volatility', 'value', 'reversion']

    def __init__(self, mean=0, volatility=1, value=None, reversion=0):
        self.value = value if value is not None else mean
        self.reversion = reversion
        self.mean = mean
        self.volatility = volatility

    def __call__(self, t=None):
        return self.value

    def __repr__(self):
        return "MeanRevertingProcess(mean=%r, volatility=%r, value=%r, reversion=%r)" % (
            self.mean, self.volatility, self.value, self.reversion)


class Option(object):
    """Option.py:17-56."""

    def __init__(self, spot=100, strike=99, interest_rate=0.06, volatility=0.2,
                 variance=None, tenor=1.0):
        self.attrs = ['spot', 'strike', 'interest_rate', 'variance', 'tenor']
        self.Type = "Option"
        self.spot = float(spot)
        self.strike = float(strike)
        if isinstance(interest_rate, MeanRevertingProcess):
            self.interest_rate = interest_rate
        else:
            self.interest_rate = MeanRevertingProcess(mean=interest_rate, volatility=0)
        if isinstance(variance, MeanRevertingProcess):
            self._variance = variance
        elif variance is not None:
            self._variance = MeanRevertingProcess(mean=variance, volatility=0)
        else:
            self._variance = MeanRevertingProcess(mean=volatility ** 2.0, volatility=0)
        self.tenor = float(tenor)

    @property
    def variance(self):
        return self._variance

    @property
    def volatility(self):
        return self.variance.value ** 0.5


class BarrierOption(Option):
    """Option.py:183-240: ``top``/``bottom`` are (knock_in, level) or None."""

    def __init__(self, spot=100, strike=99, interest_rate=0.06, volatility=0.2,
                 variance=None, tenor=1.0, top=None, bottom=None):
        Option.__init__(self, spot=spot, strike=strike, interest_rate=interest_rate,
                        volatility=volatility, variance=variance, tenor=tenor)
        self.Type = "BarrierOption"
        self.top = top
        self.bottom = bottom


class HestonOption(Option):
    """heston.py:37-79."""

    def __init__(self, spot=100, strike=99, interest_rate=0.06, volatility=0.2,
                 variance=None, tenor=1.0, mean_reversion=1, mean_variance=None,
                 vol_of_variance=None, correlation=0):
        Option.__init__(self, spot=spot, strike=strike, interest_rate=interest_rate,
                        variance=variance, volatility=volatility, tenor=tenor)
        self.Type = "HestonOption"
        self.attrs += ['correlation']
        self.variance.reversion = mean_reversion
        self.variance.mean = mean_variance if mean_variance is not None else self.variance.value
        self.variance.volatility = 0.4 if vol_of_variance is None else vol_of_variance
        self.correlation = correlation


class HestonBarrierOption(HestonOption):
    """heston.py:219-253."""

    def __init__(self, spot=100, strike=99, interest_rate=0.06, volatility=0.2,
                 variance=None, tenor=1.0, mean_reversion=1, mean_variance=None,
                 vol_of_variance=0.4, correlation=0, top=None, bottom=None):
        HestonOption.__init__(self, spot=spot, strike=strike, interest_rate=interest_rate,
                              volatility=volatility, variance=variance, tenor=tenor,
                              mean_reversion=mean_reversion, mean_variance=mean_variance,
                              vol_of_variance=vol_of_variance, correlation=correlation)
        self.Type = 'HestonBarrierOption'
        self.top = top
        self.bottom = bottom
